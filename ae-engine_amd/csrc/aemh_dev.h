// aemh_dev.h -- device-side types and per-particle math of the MLS-MPM step.
//
// Included by every HIP translation unit.  The kernels are compiled twice,
// k_fast.hip (FP32 mode) and k_det.hip (deterministic Q32.32 mode), both with
// -ffp-contract=off: per-particle arithmetic keeps the reference's rustc
// evaluation order (rustc never fuses multiply-adds).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "aemh_common.h"
#include "aemh_half.h"

namespace aemh {

// ---------------------------------------------------------------------------
// Quadratic B-spline, ae_particle/src/mpm_solver.rs:442-464.  NaN input gives
// 0 (both comparisons fail), as in the reference.
// ---------------------------------------------------------------------------
// Both branches are evaluated and selected: the same operations on the same
// operands as the branching form, without the divergent branches that the
// per-lane fractional positions would otherwise cause.
__device__ __forceinline__ float b_spline(float x) {
  const float ax = fabsf(x);
  const float inner = 0.75f - ax * ax;
  const float t = 1.5f - ax;
  const float outer = 0.5f * t * t;
  return ax < 0.5f ? inner : (ax < 1.5f ? outer : 0.0f);
}
__device__ __forceinline__ float b_spline_deriv(float x) {
  const float ax = fabsf(x);
  const float inner = -2.0f * x;
  const float sg = signbit(x) ? -1.0f : 1.0f;  // f32::signum (x is never NaN here)
  const float outer = -(1.5f - ax) * sg;
  return ax < 0.5f ? inner : (ax < 1.5f ? outer : 0.0f);
}

// ---------------------------------------------------------------------------
// Stencil base: base = floor((x - o) inv_dx - 0.5), fx = t - base, base
// clamped at 0 (mpm_solver.rs:853-861).  Written with _rn intrinsics so the
// sort key (computed in one translation unit) and P2G/G2P (another, maybe
// contracted) always agree on the base cell, and match the reference bits.
// The int conversion is clamped to +-2^30: beyond that the particle lies so
// far outside any grid that it touches no node either way.
// ---------------------------------------------------------------------------
struct Stencil {
  int g[3];     // clamped base (>= 0)
  float fx[3];  // fractional position, taken with the unclamped base
};
__device__ __forceinline__ void stencil_axis(float x, float o, float inv_dx, int& g,
                                             float& fx) {
  float t = __fsub_rn(__fmul_rn(__fsub_rn(x, o), inv_dx), 0.5f);
  float fl = floorf(t);
  fx = __fsub_rn(t, fl);
  float cl = fminf(fmaxf(fl, -1073741824.0f), 1073741824.0f);
  int gi = (int)cl;
  g = gi > 0 ? gi : 0;
}
__device__ __forceinline__ Stencil make_stencil(const SceneDev& s, float x, float y, float z) {
  Stencil st;
  stencil_axis(x, s.o[0], s.inv_dx, st.g[0], st.fx[0]);
  stencil_axis(y, s.o[1], s.inv_dx, st.g[1], st.fx[1]);
  stencil_axis(z, s.o[2], s.inv_dx, st.g[2], st.fx[2]);
  return st;
}
// Global block id of a clamped base cell (sort key); bases past the grid are
// clamped into its last block -- such particles touch no node of that axis.
__device__ __forceinline__ uint32_t block_key(const SceneDev& s, const Stencil& st) {
  int bx = min(st.g[0], s.n[0] - 1) >> 2;
  int by = min(st.g[1], s.n[1] - 1) >> 2;
  int bz = min(st.g[2], s.n[2] - 1) >> 2;
  return s.block_off + (uint32_t)((bz * s.nb[1] + by) * s.nb[0] + bx);
}
// Runs of equal keys seen by one wave (how far the working set is from block
// order), spread over kRunLanes counters on separate 64 B lines.
__device__ __forceinline__ void count_runs(SortDev& s, uint32_t wave_id, uint32_t runs) {
  atomicAdd(&s.runs[(wave_id % (uint32_t)kRunLanes) * kRunStride], runs);
}
// Node address in the block-major grid (4^3 nodes per block, x fastest).
__device__ __forceinline__ uint32_t node_addr(const SceneDev& s, int gx, int gy, int gz) {
  uint32_t b = s.block_off + (uint32_t)(((gz >> 2) * s.nb[1] + (gy >> 2)) * s.nb[0] + (gx >> 2));
  return (b << 6) | (uint32_t)(((gz & 3) << 4) | ((gy & 3) << 2) | (gx & 3));
}

// ---------------------------------------------------------------------------
// Portable ln / exp of the deterministic mode: double-precision evaluation of
// ln m = 2 atanh((m-1)/(m+1)) and a Taylor exp with 2^k scaling, rounded once
// to float.  Same operation sequence as the oracle's det_logf/det_expf, so the
// host and the GPU produce identical bits without fused operations.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double det_log_d(double x) {
  if (x != x) return x;
  if (x <= 0.0) return x == 0.0 ? -__builtin_huge_val() : __builtin_nan("");
  if (x == __builtin_huge_val()) return x;
  uint64_t b = __double_as_longlong(x);
  int e = (int)((b >> 52) & 0x7ff) - 1023;
  double m = __longlong_as_double((long long)((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL));
  if (m > 1.4142135623730951) {
    m = __dmul_rn(m, 0.5);
    e += 1;
  }
  double s = __ddiv_rn(__dsub_rn(m, 1.0), __dadd_rn(m, 1.0));
  double s2 = __dmul_rn(s, s);
  double p = 1.0 / 23.0;
  p = __dadd_rn(__dmul_rn(p, s2), 1.0 / 21.0);
  p = __dadd_rn(__dmul_rn(p, s2), 1.0 / 19.0);
  p = __dadd_rn(__dmul_rn(p, s2), 1.0 / 17.0);
  p = __dadd_rn(__dmul_rn(p, s2), 1.0 / 15.0);
  p = __dadd_rn(__dmul_rn(p, s2), 1.0 / 13.0);
  p = __dadd_rn(__dmul_rn(p, s2), 1.0 / 11.0);
  p = __dadd_rn(__dmul_rn(p, s2), 1.0 / 9.0);
  p = __dadd_rn(__dmul_rn(p, s2), 1.0 / 7.0);
  p = __dadd_rn(__dmul_rn(p, s2), 1.0 / 5.0);
  p = __dadd_rn(__dmul_rn(p, s2), 1.0 / 3.0);
  p = __dadd_rn(__dmul_rn(p, s2), 1.0);
  double r = __dmul_rn(__dmul_rn(2.0, s), p);
  double ed = (double)e;
  const double ln2hi = 6.93147180369123816490e-01, ln2lo = 1.90821492927058770002e-10;
  return __dadd_rn(__dmul_rn(ed, ln2hi), __dadd_rn(__dmul_rn(ed, ln2lo), r));
}
__device__ __forceinline__ double det_exp_d(double x) {
  if (x != x) return x;
  if (x > 709.0) return __builtin_huge_val();
  if (x < -745.0) return 0.0;
  const double ln2hi = 6.93147180369123816490e-01, ln2lo = 1.90821492927058770002e-10;
  const double invln2 = 1.44269504088896338700e+00;
  double kf = floor(__dadd_rn(__dmul_rn(x, invln2), 0.5));
  double r = __dsub_rn(__dsub_rn(x, __dmul_rn(kf, ln2hi)), __dmul_rn(kf, ln2lo));
  double p = 1.0 / 87178291200.0;
  p = __dadd_rn(__dmul_rn(p, r), 1.0 / 6227020800.0);
  p = __dadd_rn(__dmul_rn(p, r), 1.0 / 479001600.0);
  p = __dadd_rn(__dmul_rn(p, r), 1.0 / 39916800.0);
  p = __dadd_rn(__dmul_rn(p, r), 1.0 / 3628800.0);
  p = __dadd_rn(__dmul_rn(p, r), 1.0 / 362880.0);
  p = __dadd_rn(__dmul_rn(p, r), 1.0 / 40320.0);
  p = __dadd_rn(__dmul_rn(p, r), 1.0 / 5040.0);
  p = __dadd_rn(__dmul_rn(p, r), 1.0 / 720.0);
  p = __dadd_rn(__dmul_rn(p, r), 1.0 / 120.0);
  p = __dadd_rn(__dmul_rn(p, r), 1.0 / 24.0);
  p = __dadd_rn(__dmul_rn(p, r), 1.0 / 6.0);
  p = __dadd_rn(__dmul_rn(p, r), 0.5);
  p = __dadd_rn(__dmul_rn(p, r), 1.0);
  p = __dadd_rn(__dmul_rn(p, r), 1.0);
  int k = (int)kf;
  int k1 = k / 2, k2 = k - k1;
  double s1 = __longlong_as_double((long long)((uint64_t)(k1 + 1023) << 52));
  double s2 = __longlong_as_double((long long)((uint64_t)(k2 + 1023) << 52));
  return __dmul_rn(__dmul_rn(p, s1), s2);
}
__device__ __forceinline__ float det_logf(float x) { return (float)det_log_d((double)x); }
__device__ __forceinline__ float det_expf(float x) {
  if (x > 89.0f) return __builtin_huge_valf();
  if (x < -104.0f) return 0.0f;
  return (float)det_exp_d((double)x);
}

// ---------------------------------------------------------------------------
// FixedPoint Q32.32 (ae_physics/src/fixed_point.rs:22-38): from_f32 truncates
// (f64(v) 2^32) toward zero with saturation, NaN -> 0; to_f32 = raw/2^32.
// ---------------------------------------------------------------------------
__device__ __forceinline__ long long fp_from_f32(float v) {
  double d = (double)v * 4294967296.0;  // exact
  if (d != d) return 0;
  if (d >= 9223372036854775807.0) return 0x7fffffffffffffffLL;
  if (d <= -9223372036854775808.0) return (long long)0x8000000000000000ULL;
  return (long long)d;
}
__device__ __forceinline__ float fp_to_f32(long long raw) {
  return __fdiv_rn((float)raw, 4294967296.0f);
}

// Row-major 3x3 determinant in the order of mpm_solver.rs:1038-1040.
template <typename T>
__device__ __forceinline__ T det3(const T* f) {
  return f[0] * (f[4] * f[8] - f[5] * f[7]) - f[1] * (f[3] * f[8] - f[5] * f[6]) +
         f[2] * (f[3] * f[7] - f[4] * f[6]);
}

}  // namespace aemh

namespace aemh {

// Scene of a sorted particle position: scenes hold contiguous, ascending
// particle ranges, so the last scene with first <= i owns i.
__device__ __forceinline__ int scene_of(const SceneDev* __restrict__ sc, int ns, uint32_t i) {
  int lo = 0, hi = ns - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (sc[mid].first <= i) lo = mid; else hi = mid - 1;
  }
  return lo;
}
// Scene of a global block id (block ranges are ascending by scene).
__device__ __forceinline__ int scene_of_block(const SceneDev* __restrict__ sc, int ns, uint32_t b) {
  int lo = 0, hi = ns - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (sc[mid].block_off <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Wave-wide inclusive prefix sum (wave64).
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
  const int lane = __lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t t = __shfl_up(v, d, 64);
    if (lane >= d) v += t;
  }
  return v;
}

}  // namespace aemh

namespace aemh {

// Slot addressing as uniform array base + 32-bit byte offset: the compiler
// then emits global loads and stores with the base in SGPRs and one shared
// VGPR offset, instead of a 64-bit address per access.  Working sets are
// below 2^30 slots, so the offsets fit.
struct Slot {
  uint32_t b4, b2;  // byte offsets in 4-byte and 2-byte per-slot arrays
};
__device__ __forceinline__ Slot slot_of(uint32_t i) { return Slot{i << 2, i << 1}; }
template <typename T>
__device__ __forceinline__ T& at4(T* base, const Slot& s) {
  return *(T*)((char*)base + s.b4);
}
template <typename T>
__device__ __forceinline__ T& at2(T* base, const Slot& s) {
  return *(T*)((char*)base + s.b2);
}
// Element j of a 4-byte per-slot array, by index or by Slot.
template <typename T>
__device__ __forceinline__ T& el4(T* base, uint32_t j) {
  return base[j];
}
template <typename T>
__device__ __forceinline__ T& el4(T* base, const Slot& s) {
  return at4(base, s);
}

// C and F accessors honouring the storage mode (0 FP32, 1 BF16, 2 FP16).
// The 16-bit modes store F - I and pack two values per word (PBuf::CF), so
// a matrix is nine 32-bit loads or stores in every mode but FP32's 18.
__device__ __forceinline__ uint32_t enc16(float v, int storage) {
  return storage == 1 ? (uint32_t)f32_to_bf16(v) : (uint32_t)f32_to_f16(v);
}
__device__ __forceinline__ float dec_lo(uint32_t w, int storage) {
  return storage == 1 ? __uint_as_float(w << 16) : f16_to_f32((uint16_t)w);
}
__device__ __forceinline__ float dec_hi(uint32_t w, int storage) {
  return storage == 1 ? __uint_as_float(w & 0xffff0000u) : f16_to_f32((uint16_t)(w >> 16));
}
__device__ __forceinline__ uint32_t pack16(float lo, float hi, int storage) {
  return enc16(lo, storage) | (enc16(hi, storage) << 16);
}
__device__ __forceinline__ float eye3(int k) { return (k % 4 == 0) ? 1.0f : 0.0f; }

template <class I>
__device__ __forceinline__ void load_f(const PBuf& P, I j, float f[9], int storage) {
  if (storage == 0) {
#pragma unroll
    for (int k = 0; k < 9; k++) f[k] = el4(P.F[k], j);
    return;
  }
#pragma unroll
  for (int w = 0; w < 4; w++) {
    const uint32_t x = el4(P.CF[4 + w], j);
    f[2 * w] = dec_lo(x, storage) + eye3(2 * w);
    f[2 * w + 1] = dec_hi(x, storage) + eye3(2 * w + 1);
  }
  f[8] = dec_hi(el4(P.CF[8], j), storage) + 1.0f;
}
template <class I>
__device__ __forceinline__ void load_cf(const PBuf& P, I j, float c[9], float f[9], int storage) {
  if (storage == 0) {
#pragma unroll
    for (int k = 0; k < 9; k++) {
      c[k] = el4(P.C[k], j);
      f[k] = el4(P.F[k], j);
    }
    return;
  }
#pragma unroll
  for (int w = 0; w < 4; w++) {
    const uint32_t x = el4(P.CF[w], j);
    c[2 * w] = dec_lo(x, storage);
    c[2 * w + 1] = dec_hi(x, storage);
  }
  load_f(P, j, f, storage);
  c[8] = dec_lo(el4(P.CF[8], j), storage);
}
// C goes out first.  In the 16-bit modes its last component shares word 8
// with F[8], so put_c9 writes words 0-3 and put_f9 (given c[8]) the rest.
template <class I>
__device__ __forceinline__ void put_c9(PBuf& P, I j, const float c[9], int storage) {
  if (storage == 0) {
#pragma unroll
    for (int k = 0; k < 9; k++) el4(P.C[k], j) = c[k];
    return;
  }
#pragma unroll
  for (int w = 0; w < 4; w++) el4(P.CF[w], j) = pack16(c[2 * w], c[2 * w + 1], storage);
}
template <class I>
__device__ __forceinline__ void put_f9(PBuf& P, I j, const float f[9], float c8, int storage) {
  if (storage == 0) {
#pragma unroll
    for (int k = 0; k < 9; k++) el4(P.F[k], j) = f[k];
    return;
  }
#pragma unroll
  for (int w = 0; w < 4; w++)
    el4(P.CF[4 + w], j) = pack16(f[2 * w] - eye3(2 * w), f[2 * w + 1] - eye3(2 * w + 1), storage);
  el4(P.CF[8], j) = pack16(c8, f[8] - 1.0f, storage);
}
// C and F unchanged from slot p of S to slot j of D, copied in their stored
// form (no 16-bit round trip through FP32).
template <class I>
__device__ __forceinline__ void copy_cf(PBuf& D, I j, const PBuf& S, I p, int storage) {
#pragma unroll
  for (int k = 0; k < 9; k++) {
    if (storage == 0) {
      el4(D.C[k], j) = el4(S.C[k], p);
      el4(D.F[k], j) = el4(S.F[k], p);
    } else {
      el4(D.CF[k], j) = el4(S.CF[k], p);
    }
  }
}

}  // namespace aemh
