// fixedsim.cpp -- CPU restatement of the all-integer MLS-MPM solver
// MpmSimulation (ae_physics/src/mpm.rs:139-951) and of the FixedPoint
// Q32.32 arithmetic it runs on (ae_physics/src/fixed_point.rs:4-883).
//
// TEST INFRASTRUCTURE ONLY: built into oracle/libfixedsim.so and loaded by
// tests/ as the checker of the device solver (aemh_fixed_*); never by the
// product.
//
// Serial, array-of-structures, one particle after another exactly as the
// reference loops run.  Rust semantics restated in C++:
//   i64 `as` casts from f32/f64 saturate (NaN -> 0); i128 `/` truncates
//   toward zero; `as i64` of an i128 wraps; FixedPoint +, - saturate; *, /
//   are the i128 forms whose saturating fallbacks cannot trigger
//   (fixed_point.rs:63-79, 187-203, 318-348); the f32 sin/cos/tan the
//   reference calls (fracture offsets, granular friction) are glibc's, as
//   Rust's std on Linux.
//
// Pinned in tests/test_fixed_sim.py:
//   FixedPoint ops: the reference's unit tests (fixed_point.rs:889-962);
//     mul / div / sqrt / neg against Python integer arithmetic on random
//     and extreme operands; the sin table and its interpolation; ln against
//     the truncated artanh series and math.log near 1; exp against math.exp;
//     from_f32 / to_f32 against numpy; det / inverse on diagonal, random and
//     overflowing matrices.
//   MpmSimulation: the reference's unit tests (mpm.rs:957-1148); one
//     substep of free fall (P2G, grid force, G2P) in exact integers;
//     plasticity, granular and fluid returns against their float closed
//     forms; the SVD quirk (V = the rotated A^T A), U and S, the snow clamp
//     and rebuilt F in exact integers on a quarter-turned diagonal F;
//     brittle's config switch; fracture on F = 1.2 I (fragment count,
//     ring offsets, masses, ids, connections, event, no second break).
//   tests/test_oracle_pins.py: the momentum's stress term S rel / cell^2
//     (mpm.rs:322-324) with a non-symmetric S, the column layout of the
//     velocity gradient and F + (grad v F) dt (:384-419) against their
//     closed forms on one particle, and update_deformation's Kirchhoff
//     stress (:429-454) on diagonal and non-symmetric F.
// The grid scatter bound it reports is the device's quantity, compared
// exactly in the GPU tests.
//
// Planted mistakes, compiled in only by tests/test_oracle_pins.py (one
// -DORACLE_MUT_* flag per throw-away build) to show the pins fail on them:
//   ORACLE_MUT_FX_STRESS_T     momentum with S^T rel
//   ORACLE_MUT_FX_VGRAD_ROWS   velocity gradient accumulated row-wise
//   ORACLE_MUT_FX_KIRCHHOFF_F  Kirchhoff stress times F instead of F^T
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

namespace {

typedef __int128 i128;
const int64_t I64MAX = INT64_MAX, I64MIN = INT64_MIN;
const i128 SCALE = (i128)1 << 32;

struct Fx {
  int64_t raw;
};
Fx R(int64_t r) { return Fx{r}; }
const Fx ZERO{0}, ONE{(int64_t)1 << 32}, NEG_ONE{-((int64_t)1 << 32)};
const Fx FP_HALF{(int64_t)1 << 31}, FP_ONE_POINT_FIVE{6442450944LL},
    FP_THREE_QUARTERS{3221225472LL}, FP_TWO{8589934592LL}, FP_THREE{12884901888LL}, FP_EPS{10};

int64_t sat_cast_f64(double v) {  // Rust `f64 as i64`
  if (v != v) return 0;
  if (v >= 9223372036854775808.0) return I64MAX;
  if (v <= -9223372036854775808.0) return I64MIN;
  return (int64_t)v;
}
Fx from_f32(float v) { return R(sat_cast_f64((double)v * 4294967296.0)); }
Fx from_i32(int32_t v) { return R((int64_t)v * 4294967296LL); }
float to_f32(Fx a) { return (float)a.raw / 4294967296.0f; }
double to_f64(Fx a) { return (double)a.raw / 4294967296.0; }

Fx add(Fx a, Fx b) {
  int64_t r;
  if (__builtin_add_overflow(a.raw, b.raw, &r)) return R(b.raw > 0 ? I64MAX : I64MIN);
  return R(r);
}
Fx sub(Fx a, Fx b) {
  int64_t r;
  if (__builtin_sub_overflow(a.raw, b.raw, &r)) return R(b.raw < 0 ? I64MAX : I64MIN);
  return R(r);
}
Fx neg(Fx a) { return R(a.raw == I64MIN ? I64MAX : -a.raw); }
Fx fabs_(Fx a) { return R(a.raw == I64MIN ? I64MAX : (a.raw < 0 ? -a.raw : a.raw)); }
// (a * b) as i128, / 2^32 truncating, as i64 wrapping (fixed_point.rs:63-71)
Fx mul(Fx a, Fx b) {
  i128 p = (i128)a.raw * (i128)b.raw;
  return R((int64_t)(p / SCALE));
}
// (a * 2^32) / b truncating, as i64; b = 0 saturates by sign (:318-348)
Fx div(Fx a, Fx b) {
  if (b.raw == 0) return R(a.raw == 0 ? 0 : (a.raw > 0 ? I64MAX : I64MIN));
  i128 p = (i128)a.raw * SCALE;
  return R((int64_t)(p / (i128)b.raw));
}
bool lt(Fx a, Fx b) { return a.raw < b.raw; }
bool gt(Fx a, Fx b) { return a.raw > b.raw; }
Fx fclamp(Fx a, Fx lo, Fx hi) { return R(a.raw < lo.raw ? lo.raw : (a.raw > hi.raw ? hi.raw : a.raw)); }
Fx ffloor(Fx a) { return R(a.raw & ~(int64_t)0xFFFFFFFF); }

// integer_sqrt_i128 (fixed_point.rs:856-867), Newton from n
i128 isqrt(i128 n) {
  if (n <= 0) return 0;
  i128 x = n;
  i128 y = (x + 1) >> 1;
  while (y < x) {
    x = y;
    y = (x + n / x) >> 1;
  }
  return x;
}
Fx fsqrt(Fx a) {  // :81-88
  if (a.raw <= 0) return ZERO;
  return R((int64_t)isqrt((i128)a.raw * SCALE));
}
Fx fln(Fx a) {  // :254-271
  if (a.raw <= 0) return R(I64MIN);
  Fx t = div(sub(a, ONE), add(a, ONE));
  Fx t2 = mul(t, t), t3 = mul(t2, t), t5 = mul(t3, t2), t7 = mul(t5, t2), t9 = mul(t7, t2);
  Fx s = add(add(add(add(t, div(t3, from_i32(3))), div(t5, from_i32(5))), div(t7, from_i32(7))),
             div(t9, from_i32(9)));
  return mul(from_i32(2), s);
}
Fx fexp(Fx a) {  // :273-289
  if (a.raw == 0) return ONE;
  Fx result = add(ONE, a), term = a, n = from_i32(2);
  for (int it = 2; it < 12; it++) {
    term = div(mul(term, a), n);
    if (fabs_(term).raw < 1) break;
    result = add(result, term);
    n = add(n, ONE);
  }
  return result;
}

int64_t SIN_TABLE[1024];  // (sin(i / 1024 * pi/2) * 2^32) as i64 (:869-883)
void build_sin_table() {
  static bool done = false;
  if (done) return;
  const double half_pi = 1.57079632679489661923;
  for (int i = 0; i < 1024; i++) SIN_TABLE[i] = (int64_t)(sin((double)i / 1024.0 * half_pi) * 4294967296.0);
  done = true;
}
const Fx PI{13493037704LL};
Fx fsin(Fx a) {  // :110-131
  const Fx two_pi = mul(PI, from_i32(2)), half_pi = div(PI, from_i32(2));
  int64_t ang = a.raw % two_pi.raw;  // rem_euclid
  if (ang < 0) ang += two_pi.raw;
  bool negate = false;
  if (ang > PI.raw) {
    ang = two_pi.raw - ang;
    negate = true;
  }
  if (ang > half_pi.raw) ang = PI.raw - ang;
  int64_t idx = ang >> 22;
  if (idx > 1022) idx = 1022;
  const int64_t frac = ang & ((1LL << 22) - 1);
  const int64_t base = SIN_TABLE[idx], next = SIN_TABLE[idx + 1];
  const int64_t interp = base + (int64_t)((uint64_t)(next - base) * (uint64_t)frac) / (1LL << 22);
  return R(negate ? -interp : interp);
}
Fx fcos(Fx a) { return fsin(add(a, div(PI, from_i32(2)))); }
Fx atan_small(Fx x) {  // :171-180
  Fx x2 = mul(x, x), x3 = mul(x2, x), x5 = mul(x3, x2), x7 = mul(x5, x2), x9 = mul(x7, x2);
  return add(sub(add(sub(x, div(x3, from_i32(3))), div(x5, from_i32(5))), div(x7, from_i32(7))),
             div(x9, from_i32(9)));
}
Fx fatan2(Fx y, Fx x) {  // :148-169
  if (x.raw == 0 && y.raw == 0) return ZERO;
  Fx ay = fabs_(y), ax = fabs_(x), ang;
  if (ax.raw >= ay.raw) {
    ang = atan_small(div(ay, ax));
  } else {
    ang = sub(div(PI, from_i32(2)), atan_small(div(ax, ay)));
  }
  if (x.raw < 0) return y.raw >= 0 ? sub(PI, ang) : neg(sub(PI, ang));
  return y.raw >= 0 ? ang : neg(ang);
}

struct V3 {
  Fx x, y, z;
};
V3 v3(Fx x, Fx y, Fx z) { return V3{x, y, z}; }
V3 vadd(V3 a, V3 b) { return v3(add(a.x, b.x), add(a.y, b.y), add(a.z, b.z)); }
V3 vsub(V3 a, V3 b) { return v3(sub(a.x, b.x), sub(a.y, b.y), sub(a.z, b.z)); }
V3 vmul(V3 a, Fx s) { return v3(mul(a.x, s), mul(a.y, s), mul(a.z, s)); }
V3 vdiv(V3 a, Fx s) { return v3(div(a.x, s), div(a.y, s), div(a.z, s)); }
Fx dot(V3 a, V3 b) { return add(add(mul(a.x, b.x), mul(a.y, b.y)), mul(a.z, b.z)); }
Fx len2(V3 a) { return add(add(mul(a.x, a.x), mul(a.y, a.y)), mul(a.z, a.z)); }
Fx vlen(V3 a) { return fsqrt(len2(a)); }
V3 vfloor(V3 a) { return v3(ffloor(a.x), ffloor(a.y), ffloor(a.z)); }

// FixedMat3: columns x_axis, y_axis, z_axis (fixed_point.rs:626-854)
struct M3 {
  V3 c[3];
};
const M3 MZERO{{V3{ZERO, ZERO, ZERO}, V3{ZERO, ZERO, ZERO}, V3{ZERO, ZERO, ZERO}}};
const M3 MID{{V3{ONE, ZERO, ZERO}, V3{ZERO, ONE, ZERO}, V3{ZERO, ZERO, ONE}}};
Fx& el(V3& v, int r) { return r == 0 ? v.x : (r == 1 ? v.y : v.z); }
Fx elc(const V3& v, int r) { return r == 0 ? v.x : (r == 1 ? v.y : v.z); }
M3 mtrans(const M3& m) {
  M3 t;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) el(t.c[i], j) = elc(m.c[j], i);
  return t;
}
M3 madd(const M3& a, const M3& b) { return M3{{vadd(a.c[0], b.c[0]), vadd(a.c[1], b.c[1]), vadd(a.c[2], b.c[2])}}; }
M3 msub(const M3& a, const M3& b) { return M3{{vsub(a.c[0], b.c[0]), vsub(a.c[1], b.c[1]), vsub(a.c[2], b.c[2])}}; }
M3 mmuls(const M3& a, Fx s) { return M3{{vmul(a.c[0], s), vmul(a.c[1], s), vmul(a.c[2], s)}}; }
M3 mdivs(const M3& a, Fx s) { return M3{{vdiv(a.c[0], s), vdiv(a.c[1], s), vdiv(a.c[2], s)}}; }
M3 mneg(const M3& a) {
  M3 r;
  for (int i = 0; i < 3; i++) r.c[i] = v3(neg(a.c[i].x), neg(a.c[i].y), neg(a.c[i].z));
  return r;
}
M3 mmul(const M3& s, const M3& r) {  // :784-806: a = s^T, element (i,j) = a.col(i) . r.col(j)
  M3 a = mtrans(s), o;
  for (int j = 0; j < 3; j++) o.c[j] = v3(dot(a.c[0], r.c[j]), dot(a.c[1], r.c[j]), dot(a.c[2], r.c[j]));
  return o;
}
V3 mvmul(const M3& m, V3 v) {  // :808-817
  return v3(add(add(mul(m.c[0].x, v.x), mul(m.c[1].x, v.y)), mul(m.c[2].x, v.z)),
            add(add(mul(m.c[0].y, v.x), mul(m.c[1].y, v.y)), mul(m.c[2].y, v.z)),
            add(add(mul(m.c[0].z, v.x), mul(m.c[1].z, v.y)), mul(m.c[2].z, v.z)));
}
// determinant (:711-729): i128 pieces, any overflowing piece becomes 0
i128 cmul(i128 a, i128 b) {
  i128 r;
  return __builtin_mul_overflow(a, b, &r) ? 0 : r;
}
i128 csub(i128 a, i128 b) {
  i128 r;
  return __builtin_sub_overflow(a, b, &r) ? 0 : r;
}
i128 cadd(i128 a, i128 b) {
  i128 r;
  return __builtin_add_overflow(a, b, &r) ? 0 : r;
}
Fx mdet(const M3& m) {
  i128 a = m.c[0].x.raw, b = m.c[0].y.raw, c = m.c[0].z.raw, d = m.c[1].x.raw, e = m.c[1].y.raw,
       f = m.c[1].z.raw, g = m.c[2].x.raw, h = m.c[2].y.raw, i = m.c[2].z.raw;
  i128 r = cmul(a, csub(cmul(e, i), cmul(f, h)));
  r = csub(r, cmul(b, csub(cmul(d, i), cmul(f, g))));
  r = cadd(r, cmul(c, csub(cmul(d, h), cmul(e, g))));
  return R((int64_t)(r / SCALE / SCALE));
}
M3 minv(const M3& m) {  // :731-762
  Fx det = mdet(m);
  if (det.raw == 0) return MZERO;
  Fx inv = div(ONE, det);
  i128 a = m.c[0].x.raw, b = m.c[0].y.raw, c = m.c[0].z.raw, d = m.c[1].x.raw, e = m.c[1].y.raw,
       f = m.c[1].z.raw, g = m.c[2].x.raw, h = m.c[2].y.raw, i = m.c[2].z.raw;
  auto t = [&](i128 x) { return mul(R((int64_t)(x / SCALE)), inv); };
  M3 o;
  o.c[0] = v3(t(e * i - f * h), t(c * h - b * i), t(b * f - c * e));
  o.c[1] = v3(t(f * g - d * i), t(a * i - c * g), t(c * d - a * f));
  o.c[2] = v3(t(d * h - e * g), t(b * g - a * h), t(a * e - b * d));
  return o;
}

Fx bspline(Fx f) {  // mpm.rs:146-156
  Fx a = fabs_(f);
  Fx x = sub(FP_ONE_POINT_FIVE, a);
  if (a.raw < FP_HALF.raw) return sub(FP_THREE_QUARTERS, mul(f, f));
  if (a.raw < FP_ONE_POINT_FIVE.raw) return mul(mul(FP_HALF, x), x);
  return ZERO;
}

// --- SVD (mpm.rs:794-951) -------------------------------------------------
void sym_eigen(Fx a[3][3], Fx d[3]) {
  Fx v[3][3] = {{ONE, ZERO, ZERO}, {ZERO, ONE, ZERO}, {ZERO, ZERO, ONE}};
  d[0] = a[0][0];
  d[1] = a[1][1];
  d[2] = a[2][2];
  for (int it = 0; it < 50; it++) {
    Fx off = add(add(fabs_(a[0][1]), fabs_(a[0][2])), fabs_(a[1][2]));
    if (off.raw < FP_EPS.raw) break;
    for (int p = 0; p < 2; p++)
      for (int q = p + 1; q < 3; q++) {
        if (fabs_(a[p][q]).raw < FP_EPS.raw) continue;
        Fx denom = sub(d[p], d[q]);
        if (denom.raw == 0) continue;
        Fx theta = mul(FP_HALF, fatan2(div(mul(FP_TWO, a[p][q]), denom), ONE));
        Fx c = fcos(theta), s = fsin(theta);
        Fx an[3][3];
        memcpy(an, a, sizeof(an));
        an[p][p] = add(sub(mul(mul(c, c), a[p][p]), mul(mul(mul(FP_TWO, s), c), a[p][q])), mul(mul(s, s), a[q][q]));
        an[q][q] = add(add(mul(mul(s, s), a[p][p]), mul(mul(mul(FP_TWO, s), c), a[p][q])), mul(mul(c, c), a[q][q]));
        an[p][q] = ZERO;
        an[q][p] = ZERO;
        for (int r = 0; r < 3; r++)
          if (r != p && r != q) {
            an[p][r] = sub(mul(c, a[p][r]), mul(s, a[q][r]));
            an[r][p] = an[p][r];
            an[q][r] = add(mul(s, a[p][r]), mul(c, a[q][r]));
            an[r][q] = an[q][r];
          }
        memcpy(a, an, sizeof(an));
        d[p] = a[p][p];
        d[q] = a[q][q];
        for (int r = 0; r < 3; r++) {
          Fx vpr = v[r][p], vqr = v[r][q];
          v[r][p] = sub(mul(c, vpr), mul(s, vqr));
          v[r][q] = add(mul(s, vpr), mul(c, vqr));
        }
      }
  }
  // the reference returns (a, d): the rotated matrix stands for the vectors
}
void svd(const M3& m, M3& U, M3& S, M3& V) {
  Fx a[3][3];
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) a[r][c] = elc(m.c[c], r);
  Fx ata[3][3];
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) ata[i][j] = add(add(mul(a[0][i], a[0][j]), mul(a[1][i], a[1][j])), mul(a[2][i], a[2][j]));
  Fx d[3];
  sym_eigen(ata, d);  // ata now holds the rotated matrix, returned as "eigenvectors"
  Fx s[3], v[3][3], u[3][3];
  for (int i = 0; i < 3; i++) s[i] = d[i].raw > 0 ? fsqrt(d[i]) : ZERO;
  for (int i = 0; i < 3; i++)
    for (int j = 0; j < 3; j++) v[i][j] = ata[j][i];
  for (int j = 0; j < 3; j++) {
    Fx si = s[j].raw > FP_EPS.raw ? div(ONE, s[j]) : ZERO;
    for (int i = 0; i < 3; i++) {
      Fx sum = ZERO;
      for (int k = 0; k < 3; k++) sum = add(sum, mul(a[i][k], v[k][j]));
      u[i][j] = mul(sum, si);
    }
  }
  Fx det_u = add(sub(mul(u[0][0], sub(mul(u[1][1], u[2][2]), mul(u[1][2], u[2][1]))),
                     mul(u[0][1], sub(mul(u[1][0], u[2][2]), mul(u[1][2], u[2][0])))),
                 mul(u[0][2], sub(mul(u[1][0], u[2][1]), mul(u[1][1], u[2][0]))));
  if (det_u.raw < 0) {
    for (int r = 0; r < 3; r++) u[r][2] = neg(u[r][2]);
    s[2] = neg(s[2]);
  }
  for (int c = 0; c < 3; c++) U.c[c] = v3(u[0][c], u[1][c], u[2][c]);
  S = MZERO;
  S.c[0].x = s[0];
  S.c[1].y = s[1];
  S.c[2].z = s[2];
  for (int c = 0; c < 3; c++) V.c[c] = v3(v[c][0], v[c][1], v[c][2]);
}

// --- the simulation ---------------------------------------------------------
enum { ELASTIC = 0, ELASTOPLASTIC = 1, GRANULAR = 2, BRITTLE = 3, FLUID = 4, SNOW = 5 };

struct Particle {
  V3 pos{ZERO, ZERO, ZERO}, vel{ZERO, ZERO, ZERO};
  M3 F = MID;
  Fx mass = ONE, volume = ONE;
  M3 stress = MZERO;
  Fx plastic = ZERO, temperature = from_f32(293.0f);
  bool active = true;
  float color[4] = {0.8f, 0.6f, 0.4f, 1.0f};
  Fx accum = ZERO;
  bool fractured = false;
  uint64_t parent_id = 0;
  uint64_t connected[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  uint8_t num_connections = 0;
};

struct Cell {
  Fx mass = ZERO;
  V3 vel{ZERO, ZERO, ZERO};
  bool active = false;
};

struct Sim {
  uint32_t res[3];
  Fx cell, E, nu, yield, hardening, density, dt;
  V3 gravity;
  uint32_t substeps;
  int model;
  Fx model_param;
  bool enable_fracture;
  Fx fracture_strain;
  std::vector<Particle> p;
  std::vector<Cell> grid;
  double time = 0.0;
  uint64_t step_count = 0, next_id = 1;
  std::vector<double> ev_time;
  std::vector<float> ev_pos, ev_strain;
  std::vector<uint32_t> ev_frags;

  int gidx(int x, int y, int z) const {
    if (x < 0 || x >= (int)res[0] || y < 0 || y >= (int)res[1] || z < 0 || z >= (int)res[2]) return -1;
    return x + y * (int)res[0] + z * (int)res[0] * (int)res[1];
  }
  void init_particles(uint64_t count) {  // mpm.rs:179-206
    if (count == 0) return;
    Fx resy = from_i32((int32_t)res[1]);
    uint32_t ppd = (uint32_t)ceilf(cbrtf((float)count));
    Fx spacing = div(cell, from_i32((int32_t)ppd));
    for (uint32_t i = 0; i < ppd; i++)
      for (uint32_t j = 0; j < ppd; j++)
        for (uint32_t k = 0; k < ppd; k++) {
          Fx x = mul(add(from_i32((int32_t)i), FP_HALF), spacing);
          Fx y = mul(add(from_i32((int32_t)j), FP_HALF), spacing);
          Fx z = mul(add(from_i32((int32_t)k), FP_HALF), spacing);
          Particle q;
          q.pos = v3(x, add(y, mul(mul(resy, cell), from_f32(0.3f))), z);
          Fx vol = mul(mul(spacing, spacing), spacing);
          q.mass = mul(density, vol);
          q.volume = vol;
          q.parent_id = next_id++;
          p.push_back(q);
        }
  }
  static int32_t f32_to_i32(float f) {  // Rust `f32 as i32`
    if (f != f) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f <= -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
  }
  // sum over the substep's nonzero scatter contributions c of
  // (|c| >> 24) + 1, the quantity the device bounds; max over substeps
  uint64_t sum_bound = 0, max_bound = 0;
  void p2g() {
    sum_bound = 0;
    const Fx inv = div(ONE, cell);
    for (const Particle& q : p) {
      if (!q.active) continue;
      V3 s = vmul(q.pos, inv);
      V3 base = vfloor(s);
      V3 fx = vsub(vmul(q.pos, inv), base);
      const int bx = f32_to_i32(to_f32(base.x)), by = f32_to_i32(to_f32(base.y)), bz = f32_to_i32(to_f32(base.z));
      for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dz = -1; dz <= 1; dz++) {
            const int gx = bx + dx, gy = by + dy, gz = bz + dz;
            const int idx = gidx(gx, gy, gz);
            if (idx < 0) continue;
            Fx w = mul(mul(bspline(sub(fx.x, from_i32(dx))), bspline(sub(fx.y, from_i32(dy)))),
                       bspline(sub(fx.z, from_i32(dz))));
            if (w.raw < FP_EPS.raw) continue;
            V3 gp = v3(mul(from_i32(gx), cell), mul(from_i32(gy), cell), mul(from_i32(gz), cell));
            V3 rel = vsub(gp, q.pos);
#ifdef ORACLE_MUT_FX_STRESS_T
            V3 str = mvmul(mtrans(q.stress), rel);
#else
            V3 str = mvmul(q.stress, rel);
#endif
            V3 mom = vmul(vadd(q.vel, vmul(vmul(str, inv), inv)), q.mass);
            const Fx dm = mul(q.mass, w);
            const V3 dv = vmul(mom, w);
            grid[idx].mass = add(grid[idx].mass, dm);
            grid[idx].vel = vadd(grid[idx].vel, dv);
            grid[idx].active = true;
            for (int64_t c : {dm.raw, dv.x.raw, dv.y.raw, dv.z.raw})
              if (c != 0) sum_bound += ((c < 0 ? 0 - (uint64_t)c : (uint64_t)c) >> 24) + 1;
          }
    }
  }
  void grid_forces(Fx sdt) {
    if (sum_bound > max_bound) max_bound = sum_bound;
    for (Cell& c : grid) {
      if (!c.active || c.mass.raw < FP_EPS.raw) continue;
      V3 force = vmul(vmul(gravity, sdt), c.mass);
      c.vel = vadd(c.vel, force);
    }
  }
  void g2p(Fx sdt) {
    const Fx inv = div(ONE, cell);
    for (Particle& q : p) {
      if (!q.active) continue;
      V3 base = vfloor(vmul(q.pos, inv));
      V3 fx = vsub(vmul(q.pos, inv), base);
      const int bx = f32_to_i32(to_f32(base.x)), by = f32_to_i32(to_f32(base.y)), bz = f32_to_i32(to_f32(base.z));
      V3 nv{ZERO, ZERO, ZERO};
      M3 vg = MZERO;
      for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dz = -1; dz <= 1; dz++) {
            const int gx = bx + dx, gy = by + dy, gz = bz + dz;
            const int idx = gidx(gx, gy, gz);
            if (idx < 0) continue;
            Fx w = mul(mul(bspline(sub(fx.x, from_i32(dx))), bspline(sub(fx.y, from_i32(dy)))),
                       bspline(sub(fx.z, from_i32(dz))));
            if (w.raw < FP_EPS.raw || !grid[idx].active || grid[idx].mass.raw < FP_EPS.raw) continue;
            V3 gv = vdiv(grid[idx].vel, grid[idx].mass);
            nv = vadd(nv, vmul(gv, w));
            V3 gd = v3(sub(from_i32(dx), fx.x), sub(from_i32(dy), fx.y), sub(from_i32(dz), fx.z));
            for (int a = 0; a < 3; a++) {
              const Fx gva = elc(gv, a);
              V3 col = vg.c[a];
              for (int b = 0; b < 3; b++) {
                Fx contrib = mul(mul(mul(gva, w), elc(gd, b)), inv);
#ifdef ORACLE_MUT_FX_VGRAD_ROWS
                el(vg.c[b], a) = add(elc(vg.c[b], a), contrib);
                (void)col;
#else
                el(col, b) = add(elc(col, b), contrib);
#endif
              }
#ifndef ORACLE_MUT_FX_VGRAD_ROWS
              vg.c[a] = col;
#endif
            }
          }
      q.vel = nv;
      q.pos = vadd(q.pos, vmul(nv, sdt));
      q.F = madd(q.F, mmuls(mmul(vg, q.F), sdt));
      Fx det = mdet(vg);
      Fx strain = det.raw > ONE.raw ? sub(det, ONE) : sub(ONE, det);
      q.accum = add(q.accum, strain);
    }
  }
  void lame(Fx& mu, Fx& lambda) const {
    mu = div(E, mul(FP_TWO, add(ONE, nu)));
    lambda = div(mul(E, nu), mul(add(ONE, nu), sub(ONE, mul(FP_TWO, nu))));
  }
  static M3 kirchhoff_stress(const M3& f, Fx mu, Fx lambda) {
    Fx j = mdet(f);
    Fx jc = fclamp(j, from_f32(0.01f), from_f32(100.0f));
    M3 finvt = mtrans(minv(f));
    M3 pk = madd(mmuls(msub(f, finvt), mu), mmuls(finvt, mul(lambda, fln(jc))));
#ifdef ORACLE_MUT_FX_KIRCHHOFF_F
    return mdivs(mmul(pk, f), jc);
#else
    return mdivs(mmul(pk, mtrans(f)), jc);
#endif
  }
  void update_deformation() {
    Fx mu, lambda;
    lame(mu, lambda);
    for (Particle& q : p) {
      if (!q.active) continue;
      Fx j = mdet(q.F);
      if (fabs_(j).raw < FP_EPS.raw) continue;
      q.stress = kirchhoff_stress(q.F, mu, lambda);
    }
  }
  static void dev_parts(const M3& st, Fx& trace, M3& dev, Fx& norm) {
    trace = add(add(st.c[0].x, st.c[1].y), st.c[2].z);
    dev = msub(st, mmuls(MID, div(trace, FP_THREE)));
    norm = fsqrt(add(add(add(mul(dev.c[0].x, dev.c[0].x), mul(dev.c[1].y, dev.c[1].y)), mul(dev.c[2].z, dev.c[2].z)),
                     mul(FP_TWO, add(add(mul(dev.c[0].y, dev.c[0].y), mul(dev.c[0].z, dev.c[0].z)), mul(dev.c[1].z, dev.c[1].z)))));
  }
  void plasticity() {  // mpm.rs:474-502
    for (Particle& q : p) {
      if (!q.active) continue;
      Fx trace, norm;
      M3 dev;
      dev_parts(q.stress, trace, dev, norm);
      Fx vm = fsqrt(mul(FP_ONE_POINT_FIVE, norm));
      Fx limit = add(yield, mul(hardening, q.plastic));
      if (vm.raw > limit.raw) {
        Fx scale = div(limit, vm);
        q.stress = madd(mmuls(MID, div(trace, FP_THREE)), mmuls(dev, scale));
        q.plastic = add(q.plastic, div(sub(vm, limit), E));
      }
    }
  }
  void granular(Fx friction_angle) {  // :504-540
    const float rads_per_deg = 3.14159265358979323846f / 180.0f;
    const Fx mu_s = from_f32(tanf(to_f32(friction_angle) * rads_per_deg));
    for (Particle& q : p) {
      if (!q.active) continue;
      Fx trace = add(add(q.stress.c[0].x, q.stress.c[1].y), q.stress.c[2].z);
      Fx pressure = div(neg(trace), FP_THREE);
      if (pressure.raw < 0) {
        q.stress = MZERO;
        continue;
      }
      Fx tr, norm;
      M3 dev;
      dev_parts(q.stress, tr, dev, norm);
      Fx max_shear = add(mul(mu_s, pressure), yield);
      if (norm.raw > max_shear.raw) {
        Fx scale = div(max_shear, norm);
        q.stress = madd(mmuls(MID, div(trace, FP_THREE)), mmuls(dev, scale));
        q.plastic = add(q.plastic, div(sub(norm, max_shear), E));
      }
    }
  }
  void fluid(Fx viscosity) {  // :542-562
    for (Particle& q : p) {
      if (!q.active) continue;
      Fx trace = add(add(q.stress.c[0].x, q.stress.c[1].y), q.stress.c[2].z);
      Fx pressure = div(neg(trace), FP_THREE);
      if (pressure.raw < 0)
        q.stress = MZERO;
      else
        q.stress = mmuls(mneg(MID), pressure);
      Fx vm = vlen(q.vel);
      if (vm.raw > 0) {
        Fx damping = fexp(mul(neg(viscosity), vm));
        q.vel = vmul(q.vel, damping);
      }
    }
  }
  void snow() {  // :564-639
    Fx mu, lambda;
    lame(mu, lambda);
    const Fx crit_c = sub(ONE, from_f32(2.5e-2f)), crit_s = add(ONE, from_f32(7.5e-3f));
    for (Particle& q : p) {
      if (!q.active) continue;
      M3 U, S, V;
      svd(q.F, U, S, V);
      for (int i = 0; i < 3; i++) {
        Fx sv = elc(S.c[i], i);
        if (sv.raw < crit_c.raw)
          el(S.c[i], i) = crit_c;
        else if (sv.raw > crit_s.raw)
          el(S.c[i], i) = crit_s;
      }
      q.F = mmul(mmul(U, S), mtrans(V));
      q.stress = kirchhoff_stress(q.F, mu, lambda);
      Fx trace, norm;
      M3 dev;
      dev_parts(q.stress, trace, dev, norm);
      Fx vm = fsqrt(mul(FP_ONE_POINT_FIVE, norm));
      if (vm.raw > hardening.raw) {
        Fx scale = div(hardening, vm);
        q.stress = madd(mmuls(MID, div(trace, FP_THREE)), mmuls(dev, scale));
        q.plastic = add(q.plastic, div(sub(vm, hardening), E));
      }
    }
  }
  void constitutive() {  // :454-472
    switch (model) {
      case ELASTIC: break;
      case ELASTOPLASTIC: plasticity(); break;
      case GRANULAR: granular(model_param); break;
      case BRITTLE:
        plasticity();
        fracture_strain = model_param;
        enable_fracture = true;
        break;
      case FLUID: fluid(model_param); break;
      case SNOW:
        hardening = model_param;
        snow();
        break;
    }
  }
  void fracture() {  // :641-709
    const Fx threshold = fracture_strain;
    const size_t n = p.size();
    for (size_t i = 0; i < n; i++) {
      if (!p[i].active || p[i].fractured) continue;
      M3 f = p[i].F;
      M3 green = mmul(mtrans(f), f);
      Fx trace = add(add(green.c[0].x, green.c[1].y), green.c[2].z);
      Fx ps = div(trace, FP_THREE);
      Fx strain = ps.raw > ONE.raw ? sub(ps, ONE) : sub(ONE, ps);
      if (!(strain.raw > threshold.raw)) continue;
      const V3 pos = p[i].pos, vel = p[i].vel;
      const Fx mass = p[i].mass;
      float color[4];
      memcpy(color, p[i].color, sizeof(color));
      const uint64_t parent = p[i].parent_id;
      p[i].fractured = true;
      Fx ratio = div(strain, threshold);
      float rf = fminf(to_f32(ratio), 8.0f);
      uint32_t nf = (rf != rf || rf <= 0.0f) ? 0u : (uint32_t)rf;
      nf += 1;
      const Fx nffp = from_i32((int32_t)nf + 1);
      for (uint32_t fi = 0; fi < nf; fi++) {
        float angle = (float)fi * 6.28318530717958647692f / (float)nf;
        Fx half = mul(cell, from_f32(0.5f));
        V3 off = v3(mul(from_f32(cosf(angle)), half), mul(mul(from_f32(sinf(angle)), cell), from_f32(0.2f)),
                    mul(from_f32(sinf((float)fi * 1.7f)), half));
        Particle fr;
        fr.pos = vadd(pos, off);
        fr.vel = vadd(vel, vmul(off, FP_TWO));
        fr.mass = div(mass, nffp);
        fr.volume = div(p[i].volume, nffp);
        memcpy(fr.color, color, sizeof(color));
        fr.parent_id = next_id++;
        for (uint32_t c = 0; c < (nf < 7 ? nf : 7); c++) fr.connected[c] = parent + c + 1;
        fr.num_connections = (uint8_t)(nf < 8 ? nf : 8);
        p.push_back(fr);
      }
      p[i].mass = div(p[i].mass, nffp);
      ev_time.push_back(time);
      ev_pos.push_back(to_f32(pos.x));
      ev_pos.push_back(to_f32(pos.y));
      ev_pos.push_back(to_f32(pos.z));
      ev_strain.push_back(to_f32(strain));
      ev_frags.push_back(nf);
    }
  }
  void substep(Fx sdt) {
    for (Cell& c : grid) c = Cell();
    p2g();
    grid_forces(sdt);
    g2p(sdt);
    update_deformation();
    constitutive();
    if (enable_fracture) fracture();
  }
  void step() {
    Fx sdt = div(dt, from_i32((int32_t)substeps));
    for (uint32_t s = 0; s < substeps; s++) substep(sdt);
    time += to_f64(dt);
    step_count += 1;
  }
};

}  // namespace

extern "C" {

struct fx_config {
  uint32_t res[3];
  int64_t cell, E, nu, yield, hardening, density;
  int64_t gravity[3];
  int64_t dt;
  uint32_t substeps;
  uint32_t model;
  int64_t model_param;
  uint32_t enable_fracture;
  int64_t fracture_strain;
  uint64_t particle_count;
};

void* fx_sim_new(const fx_config* c) {
  build_sin_table();
  Sim* s = new Sim();
  for (int a = 0; a < 3; a++) s->res[a] = c->res[a];
  s->cell = R(c->cell);
  s->E = R(c->E);
  s->nu = R(c->nu);
  s->yield = R(c->yield);
  s->hardening = R(c->hardening);
  s->density = R(c->density);
  s->gravity = v3(R(c->gravity[0]), R(c->gravity[1]), R(c->gravity[2]));
  s->dt = R(c->dt);
  s->substeps = c->substeps;
  s->model = (int)c->model;
  s->model_param = R(c->model_param);
  s->enable_fracture = c->enable_fracture != 0;
  s->fracture_strain = R(c->fracture_strain);
  s->grid.assign((size_t)c->res[0] * c->res[1] * c->res[2], Cell());
  s->init_particles(c->particle_count);
  return s;
}
void fx_sim_free(void* h) { delete (Sim*)h; }
void fx_sim_step(void* h, uint32_t n) {
  for (uint32_t i = 0; i < n; i++) ((Sim*)h)->step();
}
uint64_t fx_sim_count(void* h) { return ((Sim*)h)->p.size(); }
uint64_t fx_sim_add_particle(void* h, const int64_t* pos, const int64_t* vel, int64_t mass,
                             const float* color) {  // mpm.rs:208-231
  Sim* s = (Sim*)h;
  Particle q;
  q.pos = v3(R(pos[0]), R(pos[1]), R(pos[2]));
  q.vel = v3(R(vel[0]), R(vel[1]), R(vel[2]));
  q.mass = R(mass);
  q.volume = div(R(mass), s->density);
  q.parent_id = s->next_id++;
  memcpy(q.color, color, sizeof(q.color));
  s->p.push_back(q);
  return q.parent_id;
}
// add_particle n times in order (default color)
uint64_t fx_sim_add_particles(void* h, uint64_t n, const int64_t* pos3, const int64_t* vel3, const int64_t* mass) {
  const float col[4] = {0.8f, 0.6f, 0.4f, 1.0f};
  uint64_t first = ((Sim*)h)->next_id;
  ((Sim*)h)->p.reserve(((Sim*)h)->p.size() + n);
  for (uint64_t i = 0; i < n; i++) fx_sim_add_particle(h, pos3 + i * 3, vel3 + i * 3, mass[i], col);
  return first;
}
int fx_sim_remove_particle(void* h, uint64_t id) {  // :233-240
  Sim* s = (Sim*)h;
  for (size_t i = 0; i < s->p.size(); i++)
    if (s->p[i].parent_id == id) {
      s->p.erase(s->p.begin() + (long)i);
      return 1;
    }
  return 0;
}
void fx_sim_set_model(void* h, uint32_t model, int64_t param) {
  ((Sim*)h)->model = (int)model;
  ((Sim*)h)->model_param = R(param);
}
void fx_sim_set_fracture(void* h, uint32_t on, int64_t threshold) {
  Sim* s = (Sim*)h;
  s->enable_fracture = on != 0;
  if (on) s->fracture_strain = R(threshold);
}
// Per particle: pos3 vel3 F9 (columns) stress9 (columns) mass volume plastic
// temperature accum = 29 int64; flags: active, fractured, num_connections;
// ids: parent_id, connected[8].
void fx_sim_read(void* h, int64_t* q29, uint8_t* flags3, uint64_t* ids9, float* color4) {
  Sim* s = (Sim*)h;
  for (size_t i = 0; i < s->p.size(); i++) {
    const Particle& q = s->p[i];
    int64_t* o = q29 + i * 29;
    o[0] = q.pos.x.raw, o[1] = q.pos.y.raw, o[2] = q.pos.z.raw;
    o[3] = q.vel.x.raw, o[4] = q.vel.y.raw, o[5] = q.vel.z.raw;
    for (int c = 0; c < 3; c++)
      for (int r = 0; r < 3; r++) {
        o[6 + c * 3 + r] = elc(q.F.c[c], r).raw;
        o[15 + c * 3 + r] = elc(q.stress.c[c], r).raw;
      }
    o[24] = q.mass.raw, o[25] = q.volume.raw, o[26] = q.plastic.raw, o[27] = q.temperature.raw;
    o[28] = q.accum.raw;
    flags3[i * 3] = q.active, flags3[i * 3 + 1] = q.fractured, flags3[i * 3 + 2] = q.num_connections;
    ids9[i * 9] = q.parent_id;
    for (int c = 0; c < 8; c++) ids9[i * 9 + 1 + c] = q.connected[c];
    memcpy(color4 + i * 4, q.color, 16);
  }
}
// assigns records [first, first+count) of the public `particles` vector
void fx_sim_write(void* h, uint64_t first, uint64_t count, const int64_t* q29, const uint8_t* flags3,
                  const uint64_t* ids9, const float* color4) {
  Sim* s = (Sim*)h;
  for (uint64_t k = 0; k < count; k++) {
    Particle& q = s->p[first + k];
    const int64_t* o = q29 + k * 29;
    q.pos = v3(R(o[0]), R(o[1]), R(o[2]));
    q.vel = v3(R(o[3]), R(o[4]), R(o[5]));
    for (int c = 0; c < 3; c++)
      for (int r = 0; r < 3; r++) {
        el(q.F.c[c], r) = R(o[6 + c * 3 + r]);
        el(q.stress.c[c], r) = R(o[15 + c * 3 + r]);
      }
    q.mass = R(o[24]), q.volume = R(o[25]), q.plastic = R(o[26]), q.temperature = R(o[27]);
    q.accum = R(o[28]);
    q.active = flags3[k * 3] != 0, q.fractured = flags3[k * 3 + 1] != 0, q.num_connections = flags3[k * 3 + 2];
    q.parent_id = ids9[k * 9];
    for (int c = 0; c < 8; c++) q.connected[c] = ids9[k * 9 + 1 + c];
    memcpy(q.color, color4 + k * 4, 16);
  }
}
void fx_sim_state(void* h, uint64_t* step_count, double* time, uint64_t* next_id, uint32_t* n_events,
                  int64_t* hardening, int64_t* fracture_strain, uint32_t* fracture_on) {
  Sim* s = (Sim*)h;
  *step_count = s->step_count;
  *time = s->time;
  *next_id = s->next_id;
  *n_events = (uint32_t)s->ev_frags.size();
  *hardening = s->hardening.raw;
  *fracture_strain = s->fracture_strain.raw;
  *fracture_on = s->enable_fracture;
}
void fx_sim_events(void* h, double* t, float* pos3, float* strain, uint32_t* frags) {
  Sim* s = (Sim*)h;
  for (size_t i = 0; i < s->ev_frags.size(); i++) {
    t[i] = s->ev_time[i];
    for (int a = 0; a < 3; a++) pos3[i * 3 + a] = s->ev_pos[i * 3 + a];
    strain[i] = s->ev_strain[i];
    frags[i] = s->ev_frags[i];
  }
}
// mpm.rs:734-765 diagnostics: total_mass, kinetic_energy, strain energy
void fx_sim_diag(void* h, int64_t* out3) {
  Sim* s = (Sim*)h;
  Fx m = ZERO, ke = ZERO, se = ZERO;
  for (const Particle& q : s->p) {
    if (!q.active) continue;
    m = add(m, q.mass);
    ke = add(ke, mul(mul(FP_HALF, q.mass), len2(q.vel)));
    se = add(se, q.accum);
  }
  out3[0] = m.raw, out3[1] = ke.raw, out3[2] = se.raw;
}
// active_particles, fractured_count, total_mass, average_temperature,
// kinetic_energy, total_strain_energy (mpm.rs:744-778), serial folds
void fx_sim_stats(void* h, int64_t* out6) {
  Sim* s = (Sim*)h;
  int64_t act = 0, frac = 0;
  Fx m = ZERO, ke = ZERO, se = ZERO, t = ZERO;
  for (const Particle& q : s->p) {
    if (q.fractured) frac++;
    if (!q.active) continue;
    act++;
    m = add(m, q.mass);
    ke = add(ke, mul(mul(FP_HALF, q.mass), len2(q.vel)));
    se = add(se, q.accum);
    t = add(t, q.temperature);
  }
  out6[0] = act, out6[1] = frac, out6[2] = m.raw;
  out6[3] = act == 0 ? from_f32(293.0f).raw : div(t, from_i32((int32_t)act)).raw;
  out6[4] = ke.raw, out6[5] = se.raw;
}
// snapshots (mpm.rs:718-742): 13 floats per active particle (position3,
// velocity3, stress magnitude, plastic strain, temperature, color4) and the
// fractured flag; returns the number written
uint64_t fx_sim_snapshots(void* h, float* out13, uint8_t* fractured) {
  Sim* s = (Sim*)h;
  uint64_t k = 0;
  for (const Particle& q : s->p) {
    if (!q.active) continue;
    const M3& st = q.stress;
    Fx mag = fsqrt(add(add(add(mul(st.c[0].x, st.c[0].x), mul(st.c[1].y, st.c[1].y)), mul(st.c[2].z, st.c[2].z)),
                       mul(FP_TWO, add(add(mul(st.c[0].y, st.c[0].y), mul(st.c[0].z, st.c[0].z)), mul(st.c[1].z, st.c[1].z)))));
    float* o = out13 + k * 13;
    o[0] = to_f32(q.pos.x), o[1] = to_f32(q.pos.y), o[2] = to_f32(q.pos.z);
    o[3] = to_f32(q.vel.x), o[4] = to_f32(q.vel.y), o[5] = to_f32(q.vel.z);
    o[6] = to_f32(mag), o[7] = to_f32(q.plastic), o[8] = to_f32(q.temperature);
    memcpy(o + 9, q.color, 16);
    fractured[k] = q.fractured;
    k++;
  }
  return k;
}
uint64_t fx_sim_scatter_bound(void* h) { return ((Sim*)h)->max_bound; }
// handle_fracture alone (mpm.rs:641-716), for its closed-form tests
void fx_sim_fracture_only(void* h) { ((Sim*)h)->fracture(); }
// handle_constitutive_model (mpm.rs:456-472) alone on one particle of the
// given F, stress (column-major), velocity and plastic strain, under the
// config's model; out: F9, stress9, velocity3, plastic, and the config
// fields the model rewrites (hardening, fracture_strain, enable_fracture)
void fx_constitutive(const fx_config* c, const int64_t* F9, const int64_t* S9, const int64_t* vel3,
                     int64_t plastic, int64_t* out25) {
  fx_config cc = *c;
  cc.particle_count = 0;
  Sim* s = (Sim*)fx_sim_new(&cc);
  Particle q;
  for (int col = 0; col < 3; col++)
    for (int r = 0; r < 3; r++) {
      el(q.F.c[col], r) = R(F9[col * 3 + r]);
      el(q.stress.c[col], r) = R(S9[col * 3 + r]);
    }
  q.vel = v3(R(vel3[0]), R(vel3[1]), R(vel3[2]));
  q.plastic = R(plastic);
  s->p.push_back(q);
  s->constitutive();
  const Particle& o = s->p[0];
  for (int col = 0; col < 3; col++)
    for (int r = 0; r < 3; r++) {
      out25[col * 3 + r] = elc(o.F.c[col], r).raw;
      out25[9 + col * 3 + r] = elc(o.stress.c[col], r).raw;
    }
  out25[18] = o.vel.x.raw, out25[19] = o.vel.y.raw, out25[20] = o.vel.z.raw;
  out25[21] = o.plastic.raw;
  out25[22] = s->hardening.raw, out25[23] = s->fracture_strain.raw, out25[24] = s->enable_fracture;
  delete s;
}
// Scalar FixedPoint operations for the unit tests.
int64_t fx_op(int op, int64_t a, int64_t b) {
  build_sin_table();
  switch (op) {
    case 0: return add(R(a), R(b)).raw;
    case 1: return sub(R(a), R(b)).raw;
    case 2: return mul(R(a), R(b)).raw;
    case 3: return div(R(a), R(b)).raw;
    case 4: return fsqrt(R(a)).raw;
    case 5: return fln(R(a)).raw;
    case 6: return fexp(R(a)).raw;
    case 7: return fsin(R(a)).raw;
    case 8: return fcos(R(a)).raw;
    case 9: return fatan2(R(a), R(b)).raw;
    case 10: return neg(R(a)).raw;
    case 11: return fabs_(R(a)).raw;
    case 12: return bspline(R(a)).raw;
  }
  return 0;
}
int64_t fx_from_f32(float v) { return from_f32(v).raw; }
float fx_to_f32(int64_t r) { return to_f32(R(r)); }
int64_t fx_sin_table(int i) {
  build_sin_table();
  return SIN_TABLE[i];
}
// Matrix operations for the unit tests: m9 columns; op 0 det (out[0]),
// 1 inverse, 2 svd -> U S V (27 values)
void fx_mat_op(int op, const int64_t* m9, int64_t* out) {
  build_sin_table();
  M3 m;
  for (int c = 0; c < 3; c++) m.c[c] = v3(R(m9[c * 3]), R(m9[c * 3 + 1]), R(m9[c * 3 + 2]));
  auto put = [&](const M3& x, int64_t* o) {
    for (int c = 0; c < 3; c++)
      for (int r = 0; r < 3; r++) o[c * 3 + r] = elc(x.c[c], r).raw;
  };
  if (op == 0) out[0] = mdet(m).raw;
  else if (op == 1) put(minv(m), out);
  else if (op == 2) {
    M3 U, S, V;
    svd(m, U, S, V);
    put(U, out);
    put(S, out + 9);
    put(V, out + 18);
  }
}

}  // extern "C"
