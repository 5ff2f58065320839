// kernels.inc.h -- P2G, grid update and G2P of one MLS-MPM substep.
//
// Compiled twice, both without multiply-add contraction (reference order):
// k_fast.hip defines AEMH_NS=fast, AEMH_DET=0 (FP32 atomics, libm ln/exp);
// k_det.hip defines AEMH_NS=det, AEMH_DET=1 (Q32.32 int64 atomics, portable
// ln/exp), which makes the whole step bit-identical to a CPU run.
//
// Reference: ae_particle/src/mpm_solver.rs
//   P2G          :831-903 (serial) / :1108-1197 (rayon fold/reduce)
//   grid update  :112-139 (+ ae_compute/src/mpm_compute.rs:343-389 margin)
//   G2P          :914-1053 (serial) / :1208-1325 (parallel)
//   velocity-only:1353-1572
//
// Skipped stencil nodes (w < 1e-10, outside the grid, grid mass < 1e-12) are
// evaluated branch-free in G2P and contribute an exact +0.0 to its sums:
// every accumulator starts at +0.0 and an IEEE sum only reaches -0.0 from two
// -0.0 operands, so adding +0.0 leaves each bit pattern unchanged -- the same
// result as the reference's `continue`.
#include "aemh_launch.h"
#include "aemh_physics.h"

namespace aemh {
namespace AEMH_NS {
namespace {

// P2G stages up to TILES (2, 3 or 4; the autotuner picks) distinct base
// blocks of one workgroup's particles in LDS; the rest scatter straight to
// HBM.  Each tile holds kTileVals int64 accumulators (6.75 KB in the FP32
// mode, 13.5 KB in the deterministic mode), so the count trades staging
// coverage against workgroups per CU: dense scenes (~300 particles per
// block) rarely need a third tile, 2 particles per cell needs three.
//
// The FP32 mode accumulates 4 fields per node: the mass and the momentum
// after the force impulse, p + f dt, with each particle's contribution
// w m (v + C d) + dt sum_c vol sigma[r][c] N'_c.  The grid update then takes
// v = (p + f dt) / m + g dt, the reference's p / m + (g + f / m) dt
// (mpm_solver.rs:121-129) with the force divided once instead of per node;
// the two differ by float rounding only.  The deterministic mode (and the
// reference-order mode's 7-plane FP32 layout) keep m, p and f separate.
#if AEMH_DET
constexpr int kTileFields = kDetFields;  // 7 sums + the magnitude bound
constexpr bool kFoldForce = false;
#else
constexpr int kTileFields = 4;           // m, (p + f dt) x 3
constexpr bool kFoldForce = true;
#endif
constexpr int kTileVals = kTileFields * kTileNodes;
// Values per tile-node record of the P2G partials (the 8 deterministic-mode
// fields, or the 4 FP32 sums).
constexpr int kPartStride = AEMH_DET ? 8 : 4;
// Lane layers of the P2G re-assignment (rank within a base cell, capped),
// and the lane counters of the pair-reduced form (kPairLayers pair layers,
// then the unpaired particles and those without a tile).
constexpr int kLayers = 16;
// Pair layers of the pair-reduced P2G (up to 8 pairs per base cell).
constexpr int kPairLayers = 8;
// Base-block tiles of grid (v, m) a G2P workgroup stages in LDS.
// G2P stages up to two 6^3 grid tiles per wave of the workgroup in LDS
// (dynamic shared memory, 3,456 B per tile), workgroups of up to 512.
constexpr int kG2PMaxWG = 512;
constexpr int kG2PMaxTiles = 16;
__host__ __device__ constexpr int g2p_tiles_for(int wg) {
  return (2 * (wg >> 6)) < kG2PMaxTiles ? 2 * (wg >> 6) : kG2PMaxTiles;
}

// Compile-time switch passed to generic lambdas.
template <bool B>
struct BoolTag {
  static constexpr bool value = B;
};

// Per-particle P2G inputs after the constitutive update.
struct P2GPart {
  float wx[3], wy[3], wz[3], dwx[3], dwy[3], dwz[3];
  float px, py, pz, vx, vy, vz, mass;
  float c[9], sv[9];
};

// The 27-node scatter; `put(f, di, dj, dk, gx, gy, gz, value)` receives each
// contribution of field f (m, p0..2, f0..2; with kFold: m and p_r + f_r, the
// force term then being computed from the dt-scaled stress the caller stored
// in sv).  Each value is the reference's expression, evaluated in its order
// (mpm_solver.rs:862-898): w = (wx wy) wz (velocity-only tier: wx (wy wz)),
// u_r = ((v_r + C_r0 dx_g) + C_r1 dy_g) + C_r2 dz_g, m w u_r = (w m) u_r,
// f_r = (sv_r0 N'_x + sv_r1 N'_y) + sv_r2 N'_z; with kFold (the FP32 mode,
// whose sums have no fixed order anyway) p_r + f_r is one fused
// multiply-add, (w m) u_r + f_r rounded once.  Their partial sums depend on
// fewer than three axes, so the loops run di, dj, dk outermost to innermost
// and each partial is formed once where its axes are fixed (a node's skip
// test would otherwise keep the compiler from sharing them); a particle adds
// one value per node, so the order in which its nodes are visited does not
// change any node's sum.
template <bool kFull, bool kFold = false, class Put>
__device__ __forceinline__ void p2g_nodes(const int n0, const int n1, const int n2, const float dx,
                                          const float o0, const float o1, const float o2,
                                          const Stencil& st, const P2GPart& q, Put&& put) {
  if (kFull) {
    // per-axis terms: C_rc d_c and sv_rc N'_c for c = y, z (x is formed in the loop)
    float cy[3][3], cz[3][3], sy[3][3], sz[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      const float dyg = ((float)(st.g[1] + a) + 0.5f) * dx + o1 - q.py;
      const float dzg = ((float)(st.g[2] + a) + 0.5f) * dx + o2 - q.pz;
#pragma unroll
      for (int r = 0; r < 3; r++) {
        cy[r][a] = q.c[3 * r + 1] * dyg;
        cz[r][a] = q.c[3 * r + 2] * dzg;
        sy[r][a] = q.sv[3 * r + 1] * q.dwy[a];
        sz[r][a] = q.sv[3 * r + 2] * q.dwz[a];
      }
    }
    const float vv[3] = {q.vx, q.vy, q.vz};
#pragma unroll
    for (int di = 0; di < 3; di++) {
      const int gx = st.g[0] + di;
      const float dxg = ((float)gx + 0.5f) * dx + o0 - q.px;
      float ux0[3], fx0[3];
#pragma unroll
      for (int r = 0; r < 3; r++) {
        ux0[r] = vv[r] + q.c[3 * r] * dxg;
        fx0[r] = q.sv[3 * r] * q.dwx[di];
      }
#pragma unroll
      for (int dj = 0; dj < 3; dj++) {
        const int gy = st.g[1] + dj;
        const float wxy = q.wx[di] * q.wy[dj];
        float uxy[3], fxy[3];
#pragma unroll
        for (int r = 0; r < 3; r++) {
          uxy[r] = ux0[r] + cy[r][dj];
          fxy[r] = fx0[r] + sy[r][dj];
        }
#pragma unroll
        for (int dk = 0; dk < 3; dk++) {
          const int gz = st.g[2] + dk;
          const float w = wxy * q.wz[dk];
          if (w < 1e-10f || gx >= n0 || gy >= n1 || gz >= n2) continue;
          const float wm = w * q.mass;
          float u[3], f[3];
#pragma unroll
          for (int r = 0; r < 3; r++) {
            u[r] = uxy[r] + cz[r][dk];
            f[r] = fxy[r] + sz[r][dk];
          }
          put(0, di, dj, dk, gx, gy, gz, wm);
          if (kFold) {
            // The FP32 mode's sums have no fixed order (the reference's own
            // >= 50,000-index path), so its momentum-plus-impulse value is
            // one fused multiply-add: a rounding step less per contribution.
            put(1, di, dj, dk, gx, gy, gz, __builtin_fmaf(wm, u[0], f[0]));
            put(2, di, dj, dk, gx, gy, gz, __builtin_fmaf(wm, u[1], f[1]));
            put(3, di, dj, dk, gx, gy, gz, __builtin_fmaf(wm, u[2], f[2]));
          } else {
            put(1, di, dj, dk, gx, gy, gz, wm * u[0]);
            put(2, di, dj, dk, gx, gy, gz, wm * u[1]);
            put(3, di, dj, dk, gx, gy, gz, wm * u[2]);
            put(4, di, dj, dk, gx, gy, gz, f[0]);
            put(5, di, dj, dk, gx, gy, gz, f[1]);
            put(6, di, dj, dk, gx, gy, gz, f[2]);
          }
        }
      }
    }
  } else {
#pragma unroll
    for (int di = 0; di < 3; di++)
#pragma unroll
      for (int dj = 0; dj < 3; dj++)
#pragma unroll
        for (int dk = 0; dk < 3; dk++) {
          const float w = q.wx[di] * (q.wy[dj] * q.wz[dk]);
          const int gx = st.g[0] + di, gy = st.g[1] + dj, gz = st.g[2] + dk;
          if (w < 1e-10f || gx >= n0 || gy >= n1 || gz >= n2) continue;
          const float wm = w * q.mass;
          put(0, di, dj, dk, gx, gy, gz, wm);
          put(1, di, dj, dk, gx, gy, gz, wm * q.vx);
          put(2, di, dj, dk, gx, gy, gz, wm * q.vy);
          put(3, di, dj, dk, gx, gy, gz, wm * q.vz);
        }
  }
}

// Value of a float from the other lane of this lane's pair (2i <-> 2i + 1):
// a DPP quad permutation [1, 0, 3, 2], every lane of the wave active.
__device__ __forceinline__ float pair_swap(float x) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xB1, 0xF, 0xF, false));
}

// Maximum of v over the wave, every lane of the wave active: DPP maxima
// within each row of 16 lanes (quad permutations [1, 0, 3, 2] and [2, 3, 0,
// 1], then the half-row and row mirrors), then the four rows' lane 0.
__device__ __forceinline__ int wave_max_i32(int v) {
  v = max(v, __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false));
  v = max(v, __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false));
  v = max(v, __builtin_amdgcn_mov_dpp(v, 0x141, 0xF, 0xF, false));
  v = max(v, __builtin_amdgcn_mov_dpp(v, 0x140, 0xF, 0xF, false));
  return max(max(__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 16)),
             max(__builtin_amdgcn_readlane(v, 32), __builtin_amdgcn_readlane(v, 48)));
}

// The 27-node scatter of a pair of lanes whose particles share a base cell
// (the FP32 mode's 4 values per node: m and p_r + f_r dt, the same values as
// p2g_nodes<true, true>).  At each node lane 2i keeps fields 0 and 2 of its
// particle and hands fields 1 and 3 to lane 2i + 1, which keeps those and
// hands back 0 and 2; `put2(di, dj, dk, a, b)` then receives the pair's sums
// of the lane's two fields (lo: m and p_y; hi: p_x and p_z).  A node that one
// particle skips (w < 1e-10, outside the grid) adds 0 from that lane, as the
// reference's `continue`; both lanes visit every node, so the lane swaps
// always have their partner active.
template <class Put2>
__device__ __forceinline__ void p2g_nodes_pair(const int n0, const int n1, const int n2, const float dx,
                                               const float o0, const float o1, const float o2,
                                               const Stencil& st, const P2GPart& q, const bool hi,
                                               Put2&& put2) {
  float cy[3][3], cz[3][3], sy[3][3], sz[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    const float dyg = ((float)(st.g[1] + a) + 0.5f) * dx + o1 - q.py;
    const float dzg = ((float)(st.g[2] + a) + 0.5f) * dx + o2 - q.pz;
#pragma unroll
    for (int r = 0; r < 3; r++) {
      cy[r][a] = q.c[3 * r + 1] * dyg;
      cz[r][a] = q.c[3 * r + 2] * dzg;
      sy[r][a] = q.sv[3 * r + 1] * q.dwy[a];
      sz[r][a] = q.sv[3 * r + 2] * q.dwz[a];
    }
  }
  const float vv[3] = {q.vx, q.vy, q.vz};
#pragma unroll
  for (int di = 0; di < 3; di++) {
    const int gx = st.g[0] + di;
    const float dxg = ((float)gx + 0.5f) * dx + o0 - q.px;
    float ux0[3], fx0[3];
#pragma unroll
    for (int r = 0; r < 3; r++) {
      ux0[r] = vv[r] + q.c[3 * r] * dxg;
      fx0[r] = q.sv[3 * r] * q.dwx[di];
    }
#pragma unroll
    for (int dj = 0; dj < 3; dj++) {
      const int gy = st.g[1] + dj;
      const float wxy = q.wx[di] * q.wy[dj];
      float uxy[3], fxy[3];
#pragma unroll
      for (int r = 0; r < 3; r++) {
        uxy[r] = ux0[r] + cy[r][dj];
        fxy[r] = fx0[r] + sy[r][dj];
      }
#pragma unroll
      for (int dk = 0; dk < 3; dk++) {
        const int gz = st.g[2] + dk;
        const float w = wxy * q.wz[dk];
        const bool keep = !(w < 1e-10f || gx >= n0 || gy >= n1 || gz >= n2);
        const float wm = w * q.mass;
        float val[4];
        val[0] = wm;
#pragma unroll
        for (int r = 0; r < 3; r++)
          val[1 + r] = __builtin_fmaf(wm, uxy[r] + cz[r][dk], fxy[r] + sz[r][dk]);
#pragma unroll
        for (int k = 0; k < 4; k++) val[k] = keep ? val[k] : 0.0f;
        const float k0 = hi ? val[1] : val[0], k1 = hi ? val[3] : val[2];
        const float s0 = hi ? val[0] : val[1], s1 = hi ? val[2] : val[3];
        put2(di, dj, dk, k0 + pair_swap(s0), k1 + pair_swap(s1));
      }
    }
  }
}

// The P2G inputs of working slot p: position, velocity, mass, the B-spline
// weights and derivatives of its stencil and, in full MPM, C and the
// volume-scaled stress of its material (material_table.get(i), else entry
// 0 as NeoHookean, mpm_solver.rs:839-849; the corrected F is discarded).
// k_p2g writes the same loads out in its body: through this function the
// compiler schedules them differently and k_p2g takes 16 B more scratch and
// 4 more VGPRs per lane.
template <int STORE, class Ln = ModeLn>
__device__ __forceinline__ void p2g_inputs(const PBuf P, uint32_t p, const SceneDev& S,
                                           const Stencil& st, const MatDev* __restrict__ mats,
                                           int nmat, bool full, P2GPart& q, Ln ln = Ln()) {
  const float inv_dx = S.inv_dx;
  q.px = P.x[0][p];
  q.py = P.x[1][p];
  q.pz = P.x[2][p];
  q.vx = P.v[0][p];
  q.vy = P.v[1][p];
  q.vz = P.v[2][p];
  q.mass = P.mass[p];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    q.wx[a] = b_spline((float)a - st.fx[0]);
    q.wy[a] = b_spline((float)a - st.fx[1]);
    q.wz[a] = b_spline((float)a - st.fx[2]);
    q.dwx[a] = b_spline_deriv((float)a - st.fx[0]) * inv_dx;
    q.dwy[a] = b_spline_deriv((float)a - st.fx[1]) * inv_dx;
    q.dwz[a] = b_spline_deriv((float)a - st.fx[2]) * inv_dx;
  }
  if (full) {
    float F[9];
    load_cf(P, p, q.c, F, STORE);
    const uint32_t mi = P.mat[p];
    const MatDev M = mats[mi < (uint32_t)nmat ? mi : (uint32_t)nmat];
    float stress[9];
    compute_stress(M, F, P.J[p], P.Jp[p], stress, ln);
    const float vol = q.mass / M.density;
#pragma unroll
    for (int k = 0; k < 9; k++) q.sv[k] = vol * stress[k];
  } else {
#pragma unroll
    for (int k = 0; k < 9; k++) q.c[k] = q.sv[k] = 0.0f;
  }
}

// trunc(|x|) as an integer for |x| < 2^63, exactly: hi = floor(|x| / 2^32)
// and lo = |x| - hi 2^32 are both exact in float (x has 24 significant bits).
__device__ __forceinline__ unsigned long long trunc_abs_u64(float ax) {
  const float hi = floorf(ax * 2.3283064365386963e-10f);
  const float lo = ax - hi * 4294967296.0f;
  return ((unsigned long long)(unsigned)hi << 32) | (unsigned long long)(unsigned)lo;
}
__device__ __forceinline__ long long apply_sign(unsigned long long m, float x) {
  return x < 0.0f ? -(long long)m : (long long)m;
}

// round(y) as an int64 for |y| < 2^51, in four VALU: y + 1.5 2^52 lies in
// [2^52, 2^53), where the double ulp is 1, so its bit pattern minus that of
// 1.5 2^52 is the rounded integer in two's complement (the constant's low
// word is zero, so only the high word is adjusted).
__device__ __forceinline__ long long rint_i64(double y) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(y + 6755399441055744.0);
  return (long long)(b - 0x4338000000000000ull);
}
// rint(v * scale) for a power-of-two scale with |v scale| < 2^51, biased by
// the bit pattern of 1.5 2^52, in two VALU: the product is exact, so one
// fused multiply-add with that constant rounds exactly as the separate
// product and sum do, and its bit pattern is 0x4338000000000000 + rint(v
// scale).  The bias is a multiple of 2^51 (0x4338 = 8 * 0x867), so a sum of
// such patterns modulo 2^64 keeps the sum of the rounded values in its low
// 51 bits: unbias_sum51 recovers it whenever its magnitude is below 2^50.
__device__ __forceinline__ unsigned long long rint_biased_scaled(float v, double scale, double magic) {
  return (unsigned long long)__double_as_longlong(__builtin_fma((double)v, scale, magic));
}
__device__ __forceinline__ long long unbias_sum51(unsigned long long raw) {
  return (long long)(raw << 13) >> 13;
}

#if AEMH_DET
// FixedPoint::from_f32: trunc(v 2^32) with saturation, NaN -> 0.  The product
// is exact in double; below 2^19 its truncation goes through rint_i64.
__device__ __forceinline__ long long to_q32(float v) {
  const float av = fabsf(v);
  if (av < 524288.0f) return rint_i64(trunc((double)v * 4294967296.0));
  if (!(av < 2147483648.0f)) return v != v ? 0 : (v > 0.0f ? 0x7fffffffffffffffLL : (long long)0x8000000000000000ULL);
  return apply_sign(trunc_abs_u64(av * 4294967296.0f), v);
}
#endif

// Node t (< 216) of a 6^3 tile, enumerated block by block over the 2x2x2
// 4^3-blocks the tile overlaps (64, 32, 32, 16, 32, 16, 16, 8 nodes), so the
// lanes of one flush wave land in few 64-node grid blocks.
__device__ __forceinline__ void tile_node_of(int t, int& lx, int& ly, int& lz) {
  const int g = (t >= 64) + (t >= 96) + (t >= 128) + (t >= 144) + (t >= 176) + (t >= 192) + (t >= 208);
  constexpr int kStart[8] = {0, 64, 96, 128, 144, 176, 192, 208};
  int i = t;
#pragma unroll
  for (int k = 1; k < 8; k++)
    if (g == k) i = t - kStart[k];
  const int bx = g & 1, by = (g >> 1) & 1, bz = g >> 2;
  const int wx = bx ? 2 : 4, wy = by ? 2 : 4;
  lx = bx * 4 + i % wx;
  ly = by * 4 + (i / wx) % wy;
  lz = bz * 4 + i / (wx * wy);
}

// P2G.  One particle per thread over the block-sorted order.  The distinct
// base blocks of the workgroup (usually one or two) each get a 6^3-node LDS
// tile of int64 accumulators -- LDS integer atomics run ~30x the rate of
// ds_add_f32 on gfx950.  The deterministic mode adds FixedPoint Q32.32 values;
// the FP32 mode adds each value scaled by a per-workgroup power of two 2^s_g
// (one per field group: mass, momentum, force) chosen so 256 contributions
// cannot leave 50 bits, i.e. an exact block-floating-point sum with
// resolution 2^-42 of the group's largest contribution.  Tiles are flushed block-
// coalesced with one HBM atomic per non-zero node value.  The LDS and HBM
// scatter paths are separate code so the LDS one stays ds_add.
template <int STORE, int TILES, bool PAIRS = false>
__global__ __launch_bounds__(256) void k_p2g(PBuf P, const SceneDev* __restrict__ sc, int ns,
                                             const MatDev* __restrict__ mats, int nmat,
                                             uint32_t n, SortDev s, GridDev G, int partials,
                                             int layers) {
  __shared__ long long tile[TILES * kTileVals];
  // pair mode: lane index of each (cell, pair layer)'s first member
  __shared__ uint8_t pbase[PAIRS ? TILES * 64 * kPairLayers : 1];
  __shared__ uint32_t skey[256];
  __shared__ uint32_t sperm[256];
  __shared__ uint32_t ccnt[TILES * 64];
  __shared__ uint32_t lcnt[kLayers];
  __shared__ int16_t sslot[256];
  __shared__ uint8_t order[256];
  __shared__ int torg[TILES][3];
  __shared__ int tdim[TILES][5];  // scene grid n[0..2], nb[0..1] of the tile
  __shared__ uint32_t tboff[TILES];
  __shared__ uint32_t wcnt[4];
#if !AEMH_DET
  __shared__ int wexp[4][2];  // per wave: block exponent of each field group
#endif

  const int t = threadIdx.x;
  const int lane = __lane_id(), wave = t >> 6;
  const uint32_t j0 = blockIdx.x * 256u;
  int slot, ntiles;
  // the lane's own particle, stencil and scene (kept when lanes keep their particles)
  uint32_t p_own = 0;
  Stencil st_own = {{0, 0, 0}, {0.f, 0.f, 0.f}};
  int sid_own = 0;
  {
    // ---- tiles: one per distinct base block, in sorted order ----
    const uint32_t j = j0 + (uint32_t)t;
    const bool valid = j < n;
    // The block key is computed from the position as the count kernel did
    // (sorting keeps a particle inside its scene's range, so scene_of(j) is
    // its scene): one dependent load round (perm, then x) instead of two
    // before the first barrier, and x is at hand for the stencil.
    uint32_t p = 0, key = 0xffffffffu;
    Stencil st = {{0, 0, 0}, {0.f, 0.f, 0.f}};
    int sid = 0;
    if (valid) {
      p = s.perm[j];
      sid = scene_of(sc, ns, j);
      st = make_stencil(sc[sid], P.x[0][p], P.x[1][p], P.x[2][p]);
      key = block_key(sc[sid], st);
    }
    p_own = p;
    st_own = st;
    sid_own = sid;
    skey[t] = key;
    sperm[t] = p;
    if (t < kLayers) lcnt[t] = 0;
    __syncthreads();
    const bool head = valid && (t == 0 || skey[t - 1] != key);
    const unsigned long long hb = __ballot(head);
    const unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    if (lane == 0) wcnt[wave] = (uint32_t)__popcll(hb);
    for (int e = t; e < TILES * 64; e += 256) ccnt[e] = 0;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < 4; w++) {
      if (w < wave) pre += wcnt[w];
      tot += wcnt[w];
    }
    slot = (int)(pre + (uint32_t)__popcll(hb & le)) - 1;
    ntiles = (int)min(tot, (uint32_t)TILES);
    if (partials && head) {
      // The workgroup holding a block's first particle records the block's
      // tile slot; later workgroups of the same block hold it in slot 0.
      if (s.off[key] + s.tsum[key / kScanTile] == j) s.bslot[key] = (uint32_t)slot;
    }
    if (head && slot < TILES) {
      const SceneDev& T = sc[sid];
#pragma unroll
      for (int a = 0; a < 3; a++) {
        torg[slot][a] = (min(st.g[a], T.n[a] - 1) >> 2) * 4;
        tdim[slot][a] = T.n[a];
      }
      tdim[slot][3] = T.nb[0];
      tdim[slot][4] = T.nb[1];
      tboff[slot] = T.block_off;
    }
    sslot[t] = (int16_t)min(slot, 32767);
    __syncthreads();

    // ---- layers (an autotuned option): rank of the particle among the
    // workgroup's particles of its base cell.  Lanes are re-assigned layer by
    // layer, so the 64 lanes of one LDS atomic hit distinct nodes (the first
    // kLayers - 1 layers).  Without them lanes keep their particles, and LDS
    // atomics of one wave that hit the same node are serialized: cheaper
    // where the values to add are few and the particles per cell moderate
    // (FP32, 1M snow P2G 85.6 -> 81.3 us, 16M 0.982 -> 0.877 ms), dearer for
    // the deterministic mode's 8 sums per node, the velocity-only tier's
    // cheap contributions and dense batched scenes. ----
    if constexpr (!PAIRS) {
      if (layers) {  // uniform
        int layer = kLayers - 1;
        if (valid && slot < TILES) {
          int c = 0;
#pragma unroll
          for (int a = 2; a >= 0; a--) c = c * 4 + min(max(st.g[a] - torg[slot][a], 0), 3);
          layer = (int)min(atomicAdd(&ccnt[slot * 64 + c], 1u), (uint32_t)(kLayers - 1));
        }
        const uint32_t rl = atomicAdd(&lcnt[layer], 1u);
        __syncthreads();
        uint32_t off = 0;
        for (int l = 0; l < layer; l++) off += lcnt[l];
        order[off + rl] = (uint8_t)t;
      } else {
        order[t] = (uint8_t)t;
      }
    } else {
      // ---- pair layers: two particles of one base cell go to lanes 2i and
      // 2i + 1, which add their node values across lanes before the LDS
      // atomics.  Pair layer L holds each cell's (L+1)-th pair, so the pairs of
      // one LDS atomic still hit distinct nodes; lanes [0, npairs) are pairs,
      // then the particles left unpaired, then those not staged in a tile.
      // A particle pairs only if its base lies inside its tile's block (a
      // base past the grid's far face is clamped into the block's cells). ----
      int cell = -1;
      uint32_t rank = 0;
      if (valid && slot < TILES) {
        int c = 0;
        bool inside = true;
#pragma unroll
        for (int a = 2; a >= 0; a--) {
          const int o = st.g[a] - torg[slot][a];
          inside = inside && o >= 0 && o <= 3;
          c = c * 4 + min(max(o, 0), 3);
        }
        if (inside) {
          cell = slot * 64 + c;
          rank = atomicAdd(&ccnt[cell], 1u);
        }
      }
      __syncthreads();
      // class 0: pair member (layer pl, member rank & 1); 1: unpaired; 2: no tile
      int cls = (valid && slot < TILES) ? 1 : 2, pl = 0;
      if (cell >= 0 && rank < (ccnt[cell] & ~1u) && rank < 2u * kPairLayers) {
        cls = 0;
        pl = (int)(rank >> 1);
      }
      uint32_t idx = 0;
      if (cls == 0 && (rank & 1u) == 0u) {
        idx = atomicAdd(&lcnt[pl], 2u);
        pbase[cell * kPairLayers + pl] = (uint8_t)idx;
      } else if (cls != 0) {
        idx = atomicAdd(&lcnt[kPairLayers + cls - 1], 1u);
      }
      __syncthreads();
      if (cls == 0 && (rank & 1u)) idx = (uint32_t)pbase[cell * kPairLayers + pl] + 1u;
      const int li = cls == 0 ? pl : kPairLayers + cls - 1;
      uint32_t off = 0;
      for (int l = 0; l < li; l++) off += lcnt[l];
      order[off + idx] = (uint8_t)t;
    }
    for (int e = t; e < ntiles * kTileVals; e += 256) tile[e] = 0;
    // The cleared tiles are next touched after the block-exponent barrier;
    // only re-assigned lanes need this one, for order[].
    if (PAIRS || layers) __syncthreads();
  }
  // This lane now carries workgroup particle u (sorted position j0 + u).
  const bool own = !PAIRS && !layers;  // uniform
  const int u = own ? t : order[t];
  const uint32_t j = j0 + (uint32_t)u;
  const bool valid = j < n;
  const uint32_t p = own ? p_own : sperm[u];
  if (!own) slot = sslot[u];

  int sid = 0;
  SceneDev S{};
  Stencil st = {{0, 0, 0}, {0.f, 0.f, 0.f}};
  if (valid) {
    if (own) {
      sid = sid_own;
      S = sc[sid];
      st = st_own;
    } else {
      sid = scene_of(sc, ns, j);
      S = sc[sid];
      st = make_stencil(S, P.x[0][p], P.x[1][p], P.x[2][p]);
    }
  }

  // ---- particle inputs and stress ----------------------------------------
  const bool full = !S.vel_only;
  P2GPart q;
  q.px = q.py = q.pz = q.vx = q.vy = q.vz = q.mass = 0.0f;
  if (valid) {
    const float inv_dx = S.inv_dx;
    q.px = P.x[0][p];
    q.py = P.x[1][p];
    q.pz = P.x[2][p];
    q.vx = P.v[0][p];
    q.vy = P.v[1][p];
    q.vz = P.v[2][p];
    q.mass = P.mass[p];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      q.wx[a] = b_spline((float)a - st.fx[0]);
      q.wy[a] = b_spline((float)a - st.fx[1]);
      q.wz[a] = b_spline((float)a - st.fx[2]);
      q.dwx[a] = b_spline_deriv((float)a - st.fx[0]) * inv_dx;
      q.dwy[a] = b_spline_deriv((float)a - st.fx[1]) * inv_dx;
      q.dwz[a] = b_spline_deriv((float)a - st.fx[2]) * inv_dx;
    }
    if (full) {
      float F[9];
      load_cf(P, p, q.c, F, STORE);
      const uint32_t mi = P.mat[p];
      const MatDev M = mats[mi < (uint32_t)nmat ? mi : (uint32_t)nmat];
      float stress[9];
      compute_stress(M, F, P.J[p], P.Jp[p], stress);
      const float vol = q.mass / M.density;
      // FP32 mode: the force term enters the momentum as its impulse f dt
#pragma unroll
      for (int k = 0; k < 9; k++) q.sv[k] = kFoldForce ? S.dt * (vol * stress[k]) : vol * stress[k];
    } else {
#pragma unroll
      for (int k = 0; k < 9; k++) q.c[k] = q.sv[k] = 0.0f;
    }
  }

#if !AEMH_DET
  // ---- block exponents: |contribution of group g| < 2^e_g over the WG ----
  // mass: w m <= m; momentum after the impulse:
  // |w m u_r + dt f_r| <= m (|v_r| + 2 dx sum_c |C_rc|) + inv_dx sum_c |dt sv_rc|
  // (|w| <= 1, |N'| <= 1, |d| < 2 dx).  Bounds are doubled for rounding; a
  // non-finite bound sends the particle down the HBM path, where float
  // atomics carry NaN/inf as the reference.
  bool finite_in = true;
  int ex[2] = {-126, -126};
  if (valid) {
    float bp = 0.0f, bf = 0.0f;
    const float vv[3] = {q.vx, q.vy, q.vz};
#pragma unroll
    for (int r = 0; r < 3; r++) {
      const float tp = fabsf(vv[r]) + 2.0f * S.dx * (fabsf(q.c[3 * r]) + fabsf(q.c[3 * r + 1]) + fabsf(q.c[3 * r + 2]));
      const float tf = S.inv_dx * (fabsf(q.sv[3 * r]) + fabsf(q.sv[3 * r + 1]) + fabsf(q.sv[3 * r + 2]));
      // fmaxf would drop a NaN term: test each term itself (false for NaN)
      if (!(tp <= 3.0e38f) || !(tf <= 3.0e38f)) finite_in = false;
      bp = fmaxf(bp, tp);
      bf = fmaxf(bf, tf);
    }
    const float b[2] = {2.0f * q.mass, 2.0f * (q.mass * bp + bf)};
#pragma unroll
    for (int g = 0; g < 2; g++) {
      if (!(b[g] <= 3.0e38f)) finite_in = false;
      int e = 0;
      frexpf(b[g], &e);
      ex[g] = b[g] > 0.0f ? e : -126;
    }
  }
#pragma unroll
  for (int g = 0; g < 2; g++) {
    const int e = wave_max_i32(finite_in ? ex[g] : -126);
    if (lane == 0) wexp[wave][g] = e;
  }
#endif
  __syncthreads();

#if !AEMH_DET
  // field f belongs to group f ? 1 : 0 (mass; momentum after the impulse)
  int eg[2];
  double scale[2];
#pragma unroll
  for (int g = 0; g < 2; g++) {
    eg[g] = max(max(wexp[0][g], wexp[1][g]), max(wexp[2][g], wexp[3][g]));
    eg[g] = max(eg[g], -70);
    // 2^(42 - e_g): |contribution scale| < 2^41, so the sum of a tile
    // node's at most 256 contributions stays below 2^49
    scale[g] = (double)__int_as_float((127 + 42 - eg[g]) << 23);
  }
#endif

#if !AEMH_DET
  // A wave of pair lanes scatters pair sums when every particle in it is a
  // finite full-MPM one (others go down the per-particle paths).
  bool pair_wave = false;
  if constexpr (PAIRS) {
    uint32_t npair = 0;
#pragma unroll
    for (int l = 0; l < kPairLayers; l++) npair += lcnt[l];
    if ((uint32_t)(wave * 64 + 64) <= npair) pair_wave = __all(finite_in && full) != 0;
  }
#endif
  if (valid) {
    const int n0 = S.n[0], n1 = S.n[1], n2 = S.n[2];
    const float dx = S.dx, o0 = S.o[0], o1 = S.o[1], o2 = S.o[2];
#if AEMH_DET
    const bool lds_path = slot < TILES;
#else
    const bool lds_path = slot < TILES && finite_in;
#endif
    if (lds_path) {
      // A base cell inside this tile's block reaches local nodes 0..5 only.
      const int tb = slot * kTileVals +
                     ((st.g[2] - torg[slot][2]) * kTileSide + (st.g[1] - torg[slot][1])) * kTileSide +
                     (st.g[0] - torg[slot][0]);
#if !AEMH_DET
      // 1.5 2^52, held in registers: as a literal, the compiler rebuilds it
      // with two moves at every node after a fused multiply-add overwrote it.
      // The empty asm emits nothing.
      double magic = 6755399441055744.0;
      asm volatile("" : "+v"(magic));
#endif
#if AEMH_DET
      unsigned long long magn[27];
#pragma unroll
      for (int k = 0; k < 27; k++) magn[k] = 0;
      auto put = [&](int f, int di, int dj, int dk, int, int, int, float v) {
        const long long a = to_q32(v);
        const unsigned long long m = a < 0 ? 0ull - (unsigned long long)a : (unsigned long long)a;
        const int k = (dk * 3 + dj) * 3 + di;
        magn[k] = m > magn[k] ? m : magn[k];
        atomicAdd((unsigned long long*)&tile[tb + f * kTileNodes + (dk * kTileSide + dj) * kTileSide + di],
                  (unsigned long long)a);
      };
#else
      auto put = [&](int f, int di, int dj, int dk, int, int, int, float v) {
        const int g = f == 0 ? 0 : 1;
        atomicAdd((unsigned long long*)&tile[tb + f * kTileNodes + (dk * kTileSide + dj) * kTileSide + di],
                  rint_biased_scaled(v, scale[g], magic));
      };
#endif
#if AEMH_DET
      if (full)
        p2g_nodes<true, kFoldForce>(n0, n1, n2, dx, o0, o1, o2, st, q, put);
      else
        p2g_nodes<false>(n0, n1, n2, dx, o0, o1, o2, st, q, put);
#else
      if (pair_wave) {
        // Lane 2i (lo) adds the pair's mass and momentum-y sums, lane 2i + 1
        // its momentum-x and momentum-z sums.
        const bool hi = (t & 1) != 0;
        const double sc0 = hi ? scale[1] : scale[0];
        const int f0 = (hi ? 1 : 0) * kTileNodes, f1 = (hi ? 3 : 2) * kTileNodes;
        auto put2 = [&](int di, int dj, int dk, float v0, float v1) {
          const int node = tb + (dk * kTileSide + dj) * kTileSide + di;
          atomicAdd((unsigned long long*)&tile[node + f0], rint_biased_scaled(v0, sc0, magic));
          atomicAdd((unsigned long long*)&tile[node + f1], rint_biased_scaled(v1, scale[1], magic));
        };
        p2g_nodes_pair(n0, n1, n2, dx, o0, o1, o2, st, q, hi, put2);
      } else if (full) {
        p2g_nodes<true, kFoldForce>(n0, n1, n2, dx, o0, o1, o2, st, q, put);
      } else {
        p2g_nodes<false>(n0, n1, n2, dx, o0, o1, o2, st, q, put);
      }
#endif
#if AEMH_DET
      // Sum-|c| bound per node in units of 2^32 raw, rounded up.
#pragma unroll
      for (int dk = 0; dk < 3; dk++)
#pragma unroll
        for (int dj = 0; dj < 3; dj++)
#pragma unroll
          for (int di = 0; di < 3; di++) {
            const unsigned long long m = magn[(dk * 3 + dj) * 3 + di];
            if (m)
              atomicAdd((unsigned long long*)&tile[tb + 7 * kTileNodes + (dk * kTileSide + dj) * kTileSide + di],
                        (m >> 32) + 1ull);
          }
#endif
    } else {
      const size_t nodes = G.nodes;
      const uint32_t boff = S.block_off;
      const int nb0 = S.nb[0], nb1 = S.nb[1];
#if AEMH_DET
      long long* const acc = G.accq;
      auto put = [&](int f, int, int, int, int gx, int gy, int gz, float v) {
        const long long a = to_q32(v);
        const uint32_t b = boff + (uint32_t)(((gz >> 2) * nb1 + (gy >> 2)) * nb0 + (gx >> 2));
        const uint32_t node = (b << 6) | (uint32_t)(((gz & 3) << 4) | ((gy & 3) << 2) | (gx & 3));
        atomicAdd((unsigned long long*)(acc + (size_t)f * nodes + node), (unsigned long long)a);
        const unsigned long long m = a < 0 ? 0ull - (unsigned long long)a : (unsigned long long)a;
        atomicAdd((unsigned long long*)(acc + (size_t)7 * nodes + node), (m >> 32) + 1ull);
      };
#else
      float* const acc = G.accf;
      auto put = [&](int f, int, int, int, int gx, int gy, int gz, float v) {
        const uint32_t b = boff + (uint32_t)(((gz >> 2) * nb1 + (gy >> 2)) * nb0 + (gx >> 2));
        const uint32_t node = (b << 6) | (uint32_t)(((gz & 3) << 4) | ((gy & 3) << 2) | (gx & 3));
        unsafeAtomicAdd(acc + (size_t)f * nodes + node, v);
      };
#endif
      if (full)
        p2g_nodes<true, kFoldForce>(n0, n1, n2, dx, o0, o1, o2, st, q, put);
      else
        p2g_nodes<false>(n0, n1, n2, dx, o0, o1, o2, st, q, put);
    }
  }
  __syncthreads();

  // ---- flush, partial mode: every tile goes to its partial slot with plain
  // stores, one kPartStride-value record per tile node (thread t < 216 owns
  // node t); the grid update adds up the partials covering each node ----
  if (partials && t < kTileNodes) {
    for (int ti = 0; ti < ntiles; ti++) {
      const size_t rec = (((size_t)blockIdx.x * TILES + ti) * kTileNodes + t) * kPartStride;
#if AEMH_DET
      longlong2* dst = reinterpret_cast<longlong2*>(G.partq + rec);
#pragma unroll
      for (int h = 0; h < kPartStride / 2; h++)
        dst[h] = make_longlong2(tile[ti * kTileVals + (2 * h) * kTileNodes + t],
                                tile[ti * kTileVals + (2 * h + 1) * kTileNodes + t]);
#else
      float v[kPartStride];
#pragma unroll
      for (int f = 0; f < kPartStride; f++) {
        const long long q =
            f < kTileFields ? unbias_sum51((unsigned long long)tile[ti * kTileVals + f * kTileNodes + t]) : 0;
        const int g = f == 0 ? 0 : 1;
        v[f] = q == 0 ? 0.0f : (float)((double)q * (double)__int_as_float((127 - 42 + eg[g]) << 23));
      }
      *reinterpret_cast<float4*>(G.partf + rec) = make_float4(v[0], v[1], v[2], v[3]);
#endif
    }
  }
  // ---- flush, atomic mode: one HBM atomic per non-zero node value; tile
  // nodes are enumerated block by block so a wave's lanes share 64-node grid
  // blocks ----
  if (!partials && t < kTileNodes) {
    const size_t nodes = G.nodes;
    int lx, ly, lz;
    tile_node_of(t, lx, ly, lz);
    const int li = (lz * kTileSide + ly) * kTileSide + lx;
    for (int ti = 0; ti < ntiles; ti++) {
      const int gx = torg[ti][0] + lx, gy = torg[ti][1] + ly, gz = torg[ti][2] + lz;
      if (gx >= tdim[ti][0] || gy >= tdim[ti][1] || gz >= tdim[ti][2]) continue;
      const uint32_t b = tboff[ti] + (uint32_t)(((gz >> 2) * tdim[ti][4] + (gy >> 2)) * tdim[ti][3] + (gx >> 2));
      const uint32_t node = (b << 6) | (uint32_t)(((gz & 3) << 4) | ((gy & 3) << 2) | (gx & 3));
#pragma unroll
      for (int f = 0; f < kTileFields; f++) {
#if AEMH_DET
        const long long v = tile[ti * kTileVals + f * kTileNodes + li];
#else
        const long long v = unbias_sum51((unsigned long long)tile[ti * kTileVals + f * kTileNodes + li]);
#endif
        if (v == 0) continue;
#if AEMH_DET
        atomicAdd((unsigned long long*)(G.accq + (size_t)f * nodes + node), (unsigned long long)v);
#else
        const int g = f == 0 ? 0 : 1;
        const float fv = (float)((double)v * (double)__int_as_float((127 - 42 + eg[g]) << 23));
        unsafeAtomicAdd(G.accf + (size_t)f * nodes + node, fv);
#endif
      }
    }
  }
}

// First sorted position of block b's particles (exclusive scan of counts).
__device__ __forceinline__ uint32_t block_begin(const SortDev& s, uint32_t b, uint32_t nb,
                                                uint32_t n) {
  return b < nb ? s.off[b] + s.tsum[b / kScanTile] : n;
}

// Grid update over the active blocks, one wave per 4^3 block (persistent).
// A node's P2G sums are its HBM-atomic part plus, in partial-flush mode, the
// partials of every tile covering it: tiles of the base blocks b - (i, j,
// k), i, j, k in {0, 1}, reach this block's nodes with local coordinate < 2
// on each shifted axis, one tile per P2G workgroup holding particles of that
// base block.  The combined sums are written back for grid readback.  NF is
// the number of accumulator planes: the deterministic mode's 8, the FP32
// mode's 4 (m and the momentum after the force impulse, the only layout
// with partials), or the reference-order mode's 7 (m, p, f).
template <int NF>
__global__ __launch_bounds__(256) void k_grid_update(const SceneDev* __restrict__ sc, int ns,
                                                     SortDev s, GridDev G, int tiles,
                                                     int partials, uint32_t n, uint32_t nb_total) {
  constexpr bool kFolded = !AEMH_DET && NF == 4;
  const uint32_t lane = __lane_id();
  const uint32_t wave = (blockIdx.x * 256u + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * 256u) >> 6;
  const uint32_t nact = s.counters[kCntActive];
  for (uint32_t a = wave; a < nact; a += nwaves) {
    const uint32_t b = s.active[a];
    const SceneDev S = sc[scene_of_block(sc, ns, b)];
    const uint32_t lb = b - S.block_off;
    const int bx = (int)(lb % (uint32_t)S.nb[0]);
    const int by = (int)((lb / (uint32_t)S.nb[0]) % (uint32_t)S.nb[1]);
    const int bz = (int)(lb / (uint32_t)(S.nb[0] * S.nb[1]));
    const int gx = bx * 4 + (int)(lane & 3), gy = by * 4 + (int)((lane >> 2) & 3),
              gz = bz * 4 + (int)(lane >> 4);
    const uint32_t node = (b << 6) | lane;
#if AEMH_DET
    long long sum[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) sum[f] = G.accq[(size_t)f * G.nodes + node];
#else
    float sum[NF];
#pragma unroll
    for (int f = 0; f < NF; f++) sum[f] = G.accf[(size_t)f * G.nodes + node];
#endif
    const int lx = (int)(lane & 3), ly = (int)((lane >> 2) & 3), lz = (int)(lane >> 4);
    // Sorted ranges and first-workgroup tile slots of the 8 candidate base
    // blocks (wave-uniform), loaded before any partial.
    uint32_t o0[8], o1[8], sl0[8];
#pragma unroll
    for (int c = 0; c < 8; c++) {
      const int i = c & 1, jj = (c >> 1) & 1, k = c >> 2;
      o0[c] = o1[c] = 0;
      sl0[c] = 0xffffffffu;
      if (!partials || bz < k || by < jj || bx < i) continue;
      const uint32_t ob = S.block_off + (uint32_t)(((bz - k) * S.nb[1] + (by - jj)) * S.nb[0] + (bx - i));
      o0[c] = block_begin(s, ob, nb_total, n);
      o1[c] = block_begin(s, ob + 1, nb_total, n);
      sl0[c] = s.bslot[ob];
    }
#pragma unroll
    for (int c = 0; c < 8; c++) {
      if (o1[c] <= o0[c]) continue;
      const int i = c & 1, jj = (c >> 1) & 1, k = c >> 2;
      const bool cover = (i == 0 || lx < 2) && (jj == 0 || ly < 2) && (k == 0 || lz < 2);
      const int li = ((lz + 4 * k) * kTileSide + (ly + 4 * jj)) * kTileSide + (lx + 4 * i);
      const uint32_t w0 = o0[c] >> 8, w1 = (o1[c] - 1) >> 8;
      for (uint32_t w = w0; w <= w1; w++) {
        const uint32_t sl = (w == w0) ? sl0[c] : 0u;
        if (sl >= (uint32_t)tiles || !cover) continue;
        const size_t rec = (((size_t)w * tiles + sl) * kTileNodes + li) * kPartStride;
#if AEMH_DET
        const longlong2* src = reinterpret_cast<const longlong2*>(G.partq + rec);
#pragma unroll
        for (int h = 0; h < kPartStride / 2; h++) {
          const longlong2 q = src[h];
          sum[2 * h] += q.x;
          sum[2 * h + 1] += q.y;
        }
#else
        if constexpr (kFolded) {
          const float4 a = *reinterpret_cast<const float4*>(G.partf + rec);
          sum[0] += a.x;
          sum[1] += a.y;
          sum[2] += a.z;
          sum[3] += a.w;
        }
#endif
      }
    }
    float acc[7] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int f = 0; f < NF; f++) {
#if AEMH_DET
      G.accq[(size_t)f * G.nodes + node] = sum[f];
      if (f < 7) acc[f] = fp_to_f32(sum[f]);
#else
      G.accf[(size_t)f * G.nodes + node] = sum[f];
      acc[f] = sum[f];
#endif
    }
#if AEMH_DET
    // sum |c| >= 2^63 raw: some order of the int64 adds could have wrapped
    // (and the saturating FixedPoint sum would depend on order).
    if ((unsigned long long)sum[7] >= (1ull << 31)) s.counters[kCntOverflow] = 1u;
#endif
    const float m = acc[0];
    float v0 = acc[1], v1 = acc[2], v2 = acc[3];
    if (m > 1e-12f) {
      v0 /= m;
      v1 /= m;
      v2 /= m;
      if constexpr (kFolded) {  // the force impulse is in the momentum already
        v0 += S.g[0] * S.dt;
        v1 += S.g[1] * S.dt;
        v2 += S.g[2] * S.dt;
      } else {
        v0 += (S.g[0] + acc[4] / m) * S.dt;
        v1 += (S.g[1] + acc[5] / m) * S.dt;
        v2 += (S.g[2] + acc[6] / m) * S.dt;
      }
      if (gy == 0) {
        v1 = fmaxf(v1, 0.0f);
        v0 *= 0.9f;
        v2 *= 0.9f;
      }
      const int mg = S.bc_margin;
      if (mg > 0) {
        const float rest = S.bc_rest;
        if (gx < mg && v0 < 0.0f) v0 = -v0 * rest;
        if (gx >= S.n[0] - mg && v0 > 0.0f) v0 = -v0 * rest;
        if (gy < mg && v1 < 0.0f) v1 = -v1 * rest;
        if (gy >= S.n[1] - mg && v1 > 0.0f) v1 = -v1 * rest;
        if (gz < mg && v2 < 0.0f) v2 = -v2 * rest;
        if (gz >= S.n[2] - mg && v2 > 0.0f) v2 = -v2 * rest;
      }
    }
    G.gv[node] = make_float4(v0, v1, v2, m);
  }
}

// Block-major node address split by axis: node(x, y, z) = ax + ay + az with
// ax = (x>>2)<<6 | (x&3), ay = ((y>>2) nb0)<<6 | (y&3)<<2,
// az = (boff + (z>>2) nb1 nb0)<<6 | (z&3)<<4 -- the in-block fields are
// disjoint bits below 64, so the sum never carries into the block part.
__device__ __forceinline__ void node_axes(uint32_t boff, int nb0, int nb1, int x, int y, int z,
                                          uint32_t& ax, uint32_t& ay, uint32_t& az) {
  ax = ((uint32_t)(x >> 2) << 6) | (uint32_t)(x & 3);
  ay = ((uint32_t)((y >> 2) * nb0) << 6) | ((uint32_t)(y & 3) << 2);
  az = ((boff + (uint32_t)((z >> 2) * nb1 * nb0)) << 6) | ((uint32_t)(z & 3) << 4);
}

// The next substep's sort count, made by G2P where each new position is
// stored (slot j of the output): the count kernel's block key of that
// position and the particle's rank among the block's particles, one atomic
// per run of equal keys in a wave, as the count kernel does for slot j at the
// start of the next substep.  A wave whose lanes are split between a
// velocity-only scene and a full one counts lane by lane.  Ranks within a
// block may come in any order (so do the count kernel's).
__device__ __forceinline__ void count_next(SortDev& s, const SceneDev& S, float x0, float x1, float x2,
                                           uint32_t j, uint32_t n) {
  const uint32_t key = block_key(S, make_stencil(S, x0, x1, x2));
  s.key[j] = key;
  const int lane = __lane_id();
  const uint32_t first = j - (uint32_t)lane;  // slot of lane 0
  const uint32_t nv = min(64u, n - first);    // slots of this wave below n
  const unsigned long long all = nv == 64u ? ~0ull : ((1ull << nv) - 1ull);
  const unsigned long long act = __ballot(1);
  if (act != all) {
    s.rank[j] = atomicAdd(&s.cnt[key], 1u);
    if (lane == __ffsll((long long)act) - 1) count_runs(s, first >> 6, (uint32_t)__popcll(act));
    return;
  }
  const uint32_t prev = __shfl_up(key, 1, 64);
  const bool head = lane == 0 || prev != key;
  const unsigned long long heads = __ballot(head);
  const unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  uint32_t old = 0;
  if (head) {
    const unsigned long long above = heads & ~le;
    const int next = above ? (__ffsll((long long)above) - 1) : (int)nv;
    old = atomicAdd(&s.cnt[key], (uint32_t)(next - lane));
  }
  const int h = 63 - __clzll((long long)(heads & le));
  s.rank[j] = __shfl(old, h, 64) + (uint32_t)(lane - h);
  if (lane == 0) count_runs(s, first >> 6, (uint32_t)__popcll(heads));
}

// G2P.  Re-sort step (inplace = 0): sorted position j advances working slot
// p = perm[j] of `cur` and writes slot j of `nxt` with the particle's mass,
// material and id, so the arrays come out ordered by base block.  In-place
// step (inplace = 1): thread j advances slot j and writes it back to slot j
// of `nxt` (`cur` itself, or a scratch copy for timing), leaving the fields
// G2P does not change; particles cross blocks rarely, so between re-sorts a
// workgroup's slots still fall in a few blocks.  With COUNT, G2P also counts
// the next substep's sort (count_next), and that substep skips the count
// kernel; the instantiation without it carries none of that code.
template <int STORE, bool COUNT>
__global__ __launch_bounds__(kG2PMaxWG) void k_g2p(PBuf cur, PBuf nxt, const SceneDev* __restrict__ sc,
                                             int ns, const MatDev* __restrict__ mats, int nmat,
                                             uint32_t n, SortDev s, GridDev G, int inplace) {
  // The grid nodes of the workgroup's first g2p_tiles_for(blockDim) base
  // blocks are staged in LDS (one 6^3 tile each, loaded in one coalesced
  // pass), so the 27-node gather reads LDS instead of waiting on three
  // dependent rounds of global loads; particles of further blocks gather
  // from global memory.  Larger workgroups load fewer tiles per particle: the
  // particles of one block are split over fewer workgroups.
  extern __shared__ float4 gtile[];
  __shared__ uint32_t skey[kG2PMaxWG];
  __shared__ int gmeta[kG2PMaxTiles][8];  // origin x, y, z; grid n x, y, z; nb0; nb1
  __shared__ uint32_t gboff[kG2PMaxTiles];
  __shared__ uint32_t gkey[kG2PMaxTiles];  // block key of each staged tile
  __shared__ uint32_t wcnt[kG2PMaxWG / 64];
  __shared__ int gnonfinite;  // some staged node carries a non-finite velocity
  const int max_tiles = g2p_tiles_for((int)blockDim.x);
  const int t = threadIdx.x;
  const int lane = __lane_id(), wave = t >> 6;
  const uint32_t j = blockIdx.x * blockDim.x + (uint32_t)t;
  const bool valid = j < n;
  const SceneDev S = sc[scene_of(sc, ns, valid ? j : 0u)];
  // Every per-particle input is loaded as soon as perm is known: the loads
  // overlap the tile staging below instead of following its barriers.  The
  // block key comes from the position (the count kernel's function of the
  // same x), not from a second dependent load.
  // An in-place step walks the slots themselves: particles are independent
  // here, and the working set has stayed nearly in block order since the
  // last re-sort, so the workgroup's slots still share a few tiles.
  const uint32_t p = valid ? (inplace ? j : s.perm[j]) : 0u;
  float px = 0.0f, py = 0.0f, pz = 0.0f, pmass = 0.0f, pjp = 1.0f;
  uint16_t mi = 0;
  uint32_t pid = 0;
  float F[9];
#pragma unroll
  for (int k = 0; k < 9; k++) F[k] = (k % 4 == 0) ? 1.0f : 0.0f;
  const Slot ps = slot_of(p);
  if (valid) {
    px = at4(cur.x[0], ps);
    py = at4(cur.x[1], ps);
    pz = at4(cur.x[2], ps);
    mi = at2(cur.mat, ps);
    pmass = at4(cur.mass, ps);
    pid = at4(cur.id, ps);
    pjp = at4(cur.Jp, ps);
    // C and F are read and written out here, not through the PBuf
    // accessors: with those, the compiler kept the plane pointers of both
    // buffers live in SGPRs and spilled them to VGPR lanes (122 v_readlane
    // in the FP32 kernel, 254 in the BF16 one), which cost 3 % of G2P time.
    if (!S.vel_only) {
      if constexpr (STORE == 0) {
#pragma unroll
        for (int k = 0; k < 9; k++) F[k] = at4(cur.F[k], ps);
      } else {
#pragma unroll
        for (int w = 0; w < 4; w++) {
          const uint32_t x = at4(cur.CF[4 + w], ps);
          F[2 * w] = dec_lo(x, STORE) + eye3(2 * w);
          F[2 * w + 1] = dec_hi(x, STORE) + eye3(2 * w + 1);
        }
        F[8] = dec_hi(at4(cur.CF[8], ps), STORE) + 1.0f;
      }
    }
  }
  const Stencil st = make_stencil(S, px, py, pz);
  const uint32_t key = valid ? block_key(S, st) : 0xffffffffu;
  skey[t] = key;
  if (t == 0) gnonfinite = 0;
  __syncthreads();
  const bool head = valid && (t == 0 || skey[t - 1] != key);
  const unsigned long long hb = __ballot(head);
  const unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  if (lane == 0) wcnt[wave] = (uint32_t)__popcll(hb);
  __syncthreads();
  uint32_t pre = 0, tot = 0;
  for (int w = 0; w < (int)(blockDim.x >> 6); w++) {
    if (w < wave) pre += wcnt[w];
    tot += wcnt[w];
  }
  const int rslot = (int)(pre + (uint32_t)__popcll(hb & le)) - 1;  // the lane's key run
  const int ntiles = (int)min(tot, (uint32_t)max_tiles);
  if (head && rslot < max_tiles) {
    const int slot = rslot;
    const uint32_t lb = key - S.block_off;
    gmeta[slot][0] = (int)(lb % (uint32_t)S.nb[0]) * 4;
    gmeta[slot][1] = (int)((lb / (uint32_t)S.nb[0]) % (uint32_t)S.nb[1]) * 4;
    gmeta[slot][2] = (int)(lb / (uint32_t)(S.nb[0] * S.nb[1])) * 4;
    gmeta[slot][3] = S.n[0];
    gmeta[slot][4] = S.n[1];
    gmeta[slot][5] = S.n[2];
    gmeta[slot][6] = S.nb[0];
    gmeta[slot][7] = S.nb[1];
    gboff[slot] = S.block_off;
    gkey[slot] = key;
  }
  __syncthreads();
  // A run past the tile budget whose block an earlier run staged uses that
  // tile: on in-place steps particles that crossed a block face split their
  // block's run, so later runs repeat staged blocks.
  int slot = rslot;
  if (valid && rslot >= max_tiles) {
    for (int i = 0; i < ntiles; i++)
      if (gkey[i] == key) {
        slot = i;
        break;
      }
  }
  // Tile node li of every staged tile; outside the grid (v, m) = 0.  A node
  // with m < 1e-12, which every gather skips, is staged with v = 0, so the
  // FP32 gather's product with it is zero without a test.
  // The node loads of four tiles are issued before any is waited for (one
  // round trip per tile otherwise lengthens every workgroup's start).
  for (int li = t; li < kTileNodes; li += (int)blockDim.x) {
    const int lx = li % kTileSide, ly = (li / kTileSide) % kTileSide,
              lz = li / (kTileSide * kTileSide);
    for (int t0 = 0; t0 < ntiles; t0 += 4) {
      float4 v[4];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int ti = t0 + q;
        v[q] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (ti < ntiles) {
          const int x = gmeta[ti][0] + lx, y = gmeta[ti][1] + ly, z = gmeta[ti][2] + lz;
          if (x < gmeta[ti][3] && y < gmeta[ti][4] && z < gmeta[ti][5]) {
            uint32_t ux, uy, uz;
            node_axes(gboff[ti], gmeta[ti][6], gmeta[ti][7], x, y, z, ux, uy, uz);
            v[q] = G.gv[ux + uy + uz];
          }
        }
      }
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int ti = t0 + q;
        if (ti < ntiles) {
          float4 w = v[q];  // (0, 0, 0, 0) outside the grid
          if (w.w < 1e-12f) w.x = w.y = w.z = 0.0f;
          if (!(isfinite(w.x) && isfinite(w.y) && isfinite(w.z))) gnonfinite = 1;
          gtile[ti * kTileNodes + li] = w;
        }
      }
    }
  }
  __syncthreads();
  if (!valid) return;
  const bool use_lds = slot < max_tiles;
  const float4* __restrict__ gvp = G.gv;
  const int n0 = S.n[0], n1 = S.n[1], n2 = S.n[2];
  const int nb0 = S.nb[0], nb1 = S.nb[1];
  const uint32_t boff = S.block_off;
  float wx[3], wy[3], wz[3];
  int cx[3], cy[3], cz[3];
  bool ix[3], iy[3], iz[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    wx[a] = b_spline((float)a - st.fx[0]);
    wy[a] = b_spline((float)a - st.fx[1]);
    wz[a] = b_spline((float)a - st.fx[2]);
    ix[a] = st.g[0] + a < n0;
    iy[a] = st.g[1] + a < n1;
    iz[a] = st.g[2] + a < n2;
    cx[a] = min(st.g[0] + a, n0 - 1);
    cy[a] = min(st.g[1] + a, n1 - 1);
    cz[a] = min(st.g[2] + a, n2 - 1);
  }
  // Node addresses: coordinates are clamped into the grid, so every load is
  // in bounds; skipped nodes are masked after the load.
  uint32_t ax[3], ay[3], az[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    uint32_t u, v;
    node_axes(boff, nb0, nb1, cx[a], 0, 0, ax[a], u, v);
    node_axes(boff, nb0, nb1, 0, cy[a], 0, u, ay[a], v);
    node_axes(boff, nb0, nb1, 0, 0, cz[a], u, v, az[a]);
  }
  // LDS tile offsets (clamped coordinates lie in the base block's 6^3 halo)
  int tx[3] = {0, 0, 0}, ty[3] = {0, 0, 0}, tz[3] = {0, 0, 0};
  if (use_lds) {
#pragma unroll
    for (int a = 0; a < 3; a++) {
      tx[a] = slot * kTileNodes + (cx[a] - gmeta[slot][0]);
      ty[a] = (cy[a] - gmeta[slot][1]) * kTileSide;
      tz[a] = (cz[a] - gmeta[slot][2]) * (kTileSide * kTileSide);
    }
  }
  // Node (v, m) from the LDS tile or from the grid.  Each gather loop below
  // is instantiated once per source (a select between the two pointers
  // would turn every load into a generic flat load).
  auto lds_node = [&](int di, int dj, int dk) -> float4 { return gtile[tz[dk] + ty[dj] + tx[di]]; };
  auto hbm_node = [&](int di, int dj, int dk) -> float4 { return gvp[az[dk] + ay[dj] + ax[di]]; };
  const float dt = S.dt;
  const Slot os = slot_of(inplace ? p : j);  // output slot
  if (!inplace) {
    at4(nxt.mass, os) = pmass;
    at2(nxt.mat, os) = mi;
    at4(nxt.id, os) = pid;
  }

  if (S.vel_only) {  // mpm_solver.rs:1465-1571
    float nv0 = 0.0f, nv1 = 0.0f, nv2 = 0.0f;
    auto gather = [&](auto node_at) {
#pragma unroll
      for (int dk = 0; dk < 3; dk++)
#pragma unroll
        for (int dj = 0; dj < 3; dj++)
#pragma unroll
          for (int di = 0; di < 3; di++) {
            const float w0 = wx[di] * (wy[dj] * wz[dk]);
            const float4 g = node_at(di, dj, dk);
            // bitwise &: every lane loads the whole node (no load behind a branch)
            const bool keep = !(w0 < 1e-10f) & ix[di] & iy[dj] & iz[dk] & !(g.w < 1e-12f);
            const float w = keep ? w0 : 0.0f;
            nv0 += w * (keep ? g.x : 0.0f);
            nv1 += w * (keep ? g.y : 0.0f);
            nv2 += w * (keep ? g.z : 0.0f);
          }
    };
    if (use_lds)
      gather(lds_node);
    else
      gather(hbm_node);
    if (!isfinite(nv0) || !isfinite(nv1) || !isfinite(nv2)) nv0 = nv1 = nv2 = 0.0f;
    nv0 = fminf(fmaxf(nv0, -50.0f), 50.0f);
    nv1 = fminf(fmaxf(nv1, -50.0f), 50.0f);
    nv2 = fminf(fmaxf(nv2, -50.0f), 50.0f);
    float x0 = px + nv0 * dt, x1 = py + nv1 * dt, x2 = pz + nv2 * dt;
    if (!isfinite(x0) || !isfinite(x1) || !isfinite(x2)) {
      x0 = px;
      x1 = py;
      x2 = pz;
      nv0 = nv1 = nv2 = 0.0f;
    }
    if (x1 < 0.0f) {
      x1 = 0.0f;
      nv1 = fmaxf(nv1, 0.0f) * 0.5f;
    }
    if (x0 < 0.0f) {
      x0 = 0.0f;
      nv0 = fmaxf(nv0, 0.0f) * 0.5f;
    } else if (x0 > S.gmax_x + S.o[0]) {
      x0 = S.gmax_x + S.o[0];
      nv0 = fminf(nv0, 0.0f) * 0.5f;
    }
    if (x2 < 0.0f) {
      x2 = 0.0f;
      nv2 = fmaxf(nv2, 0.0f) * 0.5f;
    } else if (x2 > S.gmax_z + S.o[2]) {
      x2 = S.gmax_z + S.o[2];
      nv2 = fminf(nv2, 0.0f) * 0.5f;
    }
    at4(nxt.x[0], os) = x0;
    at4(nxt.x[1], os) = x1;
    at4(nxt.x[2], os) = x2;
    if constexpr (COUNT) count_next(s, S, x0, x1, x2, j, n);
    at4(nxt.v[0], os) = nv0 * S.damping;
    at4(nxt.v[1], os) = nv1 * S.damping;
    at4(nxt.v[2], os) = nv2 * S.damping;
    if (!inplace) {  // C, F, J, J_p are unchanged by this tier
      copy_cf(nxt, os, cur, ps, STORE);
      at4(nxt.J, os) = at4(cur.J, ps);
      at4(nxt.Jp, os) = pjp;
    }
    return;
  }

  float dwx[3], dwy[3], dwz[3], dxa[3], dya[3], dza[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    dwx[a] = b_spline_deriv((float)a - st.fx[0]) * S.inv_dx;
    dwy[a] = b_spline_deriv((float)a - st.fx[1]) * S.inv_dx;
    dwz[a] = b_spline_deriv((float)a - st.fx[2]) * S.inv_dx;
    dxa[a] = ((float)(st.g[0] + a) + 0.5f) * S.dx + S.o[0] - px;
    dya[a] = ((float)(st.g[1] + a) + 0.5f) * S.dx + S.o[1] - py;
    dza[a] = ((float)(st.g[2] + a) + 0.5f) * S.dx + S.o[2] - pz;
  }
  const float k4 = S.four_inv_dx2;
  float nv[3] = {0.0f, 0.0f, 0.0f};
  float nc[9] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  float gv[9] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#if AEMH_DET
  // Reference order (mpm_solver.rs:1229-1270): per node, v += w g,
  // C += (4/dx^2 w) g d^T, grad v -= w g dw^T.
  // One z-layer (9 node loads) in flight: ~133 VGPRs, 3 waves per SIMD.  The
  // fully unrolled 27-node gather needs 256 VGPRs and runs 1.6x slower.
  auto gather = [&](auto node_at) {
#pragma unroll 1
  for (int dk = 0; dk < 3; dk++)
#pragma unroll
    for (int dj = 0; dj < 3; dj++)
#pragma unroll
      for (int di = 0; di < 3; di++) {
        const float w0 = wx[di] * wy[dj] * wz[dk];
        const float4 g = node_at(di, dj, dk);
        // bitwise &: every lane loads the whole node (no load behind a branch)
        const bool keep = !(w0 < 1e-10f) & ix[di] & iy[dj] & iz[dk] & !(g.w < 1e-12f);
        const float w = keep ? w0 : 0.0f;
        const float gg[3] = {keep ? g.x : 0.0f, keep ? g.y : 0.0f, keep ? g.z : 0.0f};
        nv[0] += w * gg[0];
        nv[1] += w * gg[1];
        nv[2] += w * gg[2];
        const float scale = k4 * w;
        const float d[3] = {dxa[di], dya[dj], dza[dk]};
        const float dw[3] = {dwx[di], dwy[dj], dwz[dk]};
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int cc = 0; cc < 3; cc++) nc[3 * r + cc] += scale * gg[r] * d[cc];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int cc = 0; cc < 3; cc++) gv[3 * r + cc] -= w * gg[r] * dw[cc];
      }
  };
  if (use_lds)
    gather(lds_node);
  else
    gather(hbm_node);
#else
  // FP32 mode: the same sums regrouped by axis.  With u_r = w g_r per node,
  // C[r][c] = 4/dx^2 sum_a d_c[a] S_c[a][r] and grad v[r][c] = -sum_a dw_c[a]
  // S_c[a][r], where S_c[a][r] sums u_r over the nodes whose axis-c offset is
  // a: 12 operations per node instead of 36, the float rounding that of a
  // reassociated sum.
  float sx[3][3], sy[3][3], zc[3] = {0.0f, 0.0f, 0.0f}, zg[3] = {0.0f, 0.0f, 0.0f};
  float wxy[3][3];
#pragma unroll
  for (int a = 0; a < 3; a++)
#pragma unroll
    for (int r = 0; r < 3; r++) {
      sx[a][r] = sy[a][r] = 0.0f;
      wxy[a][r] = wx[r] * wy[a];  // [dj][di]
    }
  auto layer_sums = [&](int dk, const float t[3]) {
    const float dzk = dza[dk], dwzk = dwz[dk];
#pragma unroll
    for (int r = 0; r < 3; r++) {
      nv[r] += t[r];
      zc[r] += t[r] * dzk;
      zg[r] += t[r] * dwzk;
    }
  };
  // Staged tile: node (di, dj, dk) sits at a constant offset from the
  // particle's base node.  Nodes past the grid's far faces get weight 0 (an
  // axis weight, so w0 = 0 < 1e-10 skips them as the in-grid test does);
  // nodes with m < 1e-12 were staged with v = 0.
  float wxm[3], wym[3], wzm[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    wxm[a] = ix[a] ? wx[a] : 0.0f;
    wym[a] = iy[a] ? wy[a] : 0.0f;
    wzm[a] = iz[a] ? wz[a] : 0.0f;
  }
  // The weight test decides nothing for a particle whose node weights are
  // all either 0 or >= 1e-10: a zero weight times a finite staged velocity
  // adds +-0, which leaves every sum's bits as the reference's `continue`
  // does (the accumulators start at +0).  The smallest non-zero w0 =
  // (wx wy) wz is the product of each axis's smallest non-zero weight
  // (rounded products are monotone), so one test per particle covers its 27
  // nodes; a wave takes the test-free loop when all its LDS-path lanes pass
  // and no staged node is non-finite.
  bool tiny = false;
  if (use_lds) {
    constexpr float kBig = 3.0e38f;
    float mw[3];
    const float* ws[3] = {wxm, wym, wzm};
#pragma unroll
    for (int c = 0; c < 3; c++)
      mw[c] = fminf(fminf(ws[c][0] > 0.0f ? ws[c][0] : kBig, ws[c][1] > 0.0f ? ws[c][1] : kBig),
                    ws[c][2] > 0.0f ? ws[c][2] : kBig);
    tiny = !((mw[0] * mw[1]) * mw[2] >= 1e-10f);
  }
  const bool test_free = !__any(tiny) && !gnonfinite;
  if (use_lds) {
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
      for (int r = 0; r < 3; r++) wxy[a][r] = wxm[r] * wym[a];
    // bases past a far face (keyed into the last block) have zero weights
    // on that axis; their tile offset is clamped into the tile
    const int ox = min(st.g[0] - gmeta[slot][0], 3), oy = min(st.g[1] - gmeta[slot][1], 3),
              oz = min(st.g[2] - gmeta[slot][2], 3);
    const float4* tb = gtile + slot * kTileNodes + ox + kTileSide * oy + kTileSide * kTileSide * oz;
    // Per row of 3 nodes (fixed dj, dk) the row sum feeds the y-axis sums
    // and the layer sum once, instead of every node adding to all three.
    // With the weight test, a skipped node's u is a selected +0 (not a
    // product with a zero weight, which a non-finite grid velocity would
    // turn into NaN), the same bits as the reference skipping it.
    auto lds_gather = [&](auto tag) {
      constexpr bool kTest = decltype(tag)::value;
#pragma unroll 1
      for (int dk = 0; dk < 3; dk++) {
        const float wzk = wzm[dk];
        const float4* tk = tb + dk * (kTileSide * kTileSide);
        float t[3];
#pragma unroll
        for (int dj = 0; dj < 3; dj++) {
          float row[3];
#pragma unroll
          for (int di = 0; di < 3; di++) {
            const float w0 = wxy[dj][di] * wzk;
            const float4 g = tk[di + kTileSide * dj];
            float u[3] = {w0 * g.x, w0 * g.y, w0 * g.z};
            if (kTest) {
              const bool keep = !(w0 < 1e-10f);
#pragma unroll
              for (int r = 0; r < 3; r++) u[r] = keep ? u[r] : 0.0f;
            }
#pragma unroll
            for (int r = 0; r < 3; r++) {
              sx[di][r] += u[r];
              row[r] = di == 0 ? u[r] : row[r] + u[r];
            }
          }
#pragma unroll
          for (int r = 0; r < 3; r++) {
            sy[dj][r] += row[r];
            t[r] = dj == 0 ? row[r] : t[r] + row[r];
          }
        }
        layer_sums(dk, t);
      }
    };
    if (test_free)
      lds_gather(BoolTag<false>{});
    else
      lds_gather(BoolTag<true>{});
  } else {
#pragma unroll 1
    for (int dk = 0; dk < 3; dk++) {
      const float wzk = wz[dk];
      float t[3] = {0.0f, 0.0f, 0.0f};
#pragma unroll
      for (int dj = 0; dj < 3; dj++)
#pragma unroll
        for (int di = 0; di < 3; di++) {
          const float w0 = wxy[dj][di] * wzk;
          const float4 g = hbm_node(di, dj, dk);
          // bitwise &: every lane loads the whole node (no load behind a branch)
          const bool keep = !(w0 < 1e-10f) & ix[di] & iy[dj] & iz[dk] & !(g.w < 1e-12f);
          const float u[3] = {keep ? w0 * g.x : 0.0f, keep ? w0 * g.y : 0.0f,
                              keep ? w0 * g.z : 0.0f};
#pragma unroll
          for (int r = 0; r < 3; r++) {
            sx[di][r] += u[r];
            sy[dj][r] += u[r];
            t[r] += u[r];
          }
        }
      layer_sums(dk, t);
    }
  }
#pragma unroll
  for (int r = 0; r < 3; r++) {
    nc[3 * r + 0] = k4 * (sx[0][r] * dxa[0] + sx[1][r] * dxa[1] + sx[2][r] * dxa[2]);
    nc[3 * r + 1] = k4 * (sy[0][r] * dya[0] + sy[1][r] * dya[1] + sy[2][r] * dya[2]);
    nc[3 * r + 2] = k4 * zc[r];
    gv[3 * r + 0] = -(sx[0][r] * dwx[0] + sx[1][r] * dwx[1] + sx[2][r] * dwx[2]);
    gv[3 * r + 1] = -(sy[0][r] * dwy[0] + sy[1][r] * dwy[1] + sy[2][r] * dwy[2]);
    gv[3 * r + 2] = -zg[r];
  }
#endif
  // Q7: non-finite velocity zeroes v and C (grad v kept); +-50 clamp.
  if (!isfinite(nv[0]) || !isfinite(nv[1]) || !isfinite(nv[2])) {
    nv[0] = nv[1] = nv[2] = 0.0f;
#pragma unroll
    for (int k = 0; k < 9; k++) nc[k] = 0.0f;
  }
#pragma unroll
  for (int a = 0; a < 3; a++) {
    if (nv[a] < -50.0f) nv[a] = -50.0f;
    if (nv[a] > 50.0f) nv[a] = 50.0f;
  }
  // Q8: x += v dt with the undamped v; guards and clamps.
  float np[3] = {px + nv[0] * dt, py + nv[1] * dt, pz + nv[2] * dt};
  if (!isfinite(np[0]) || !isfinite(np[1]) || !isfinite(np[2])) {
    np[0] = 6.0f;
    np[1] = 5.0f;
    np[2] = 6.0f;
    nv[0] = nv[1] = nv[2] = 0.0f;
  }
  if (np[1] < 0.0f) {
    np[1] = 0.0f;
    nv[1] = fmaxf(nv[1], 0.0f) * 0.5f;
  }
  const float lox = S.window_clamp ? S.o[0] : 0.0f;
  const float loz = S.window_clamp ? S.o[2] : 0.0f;
  const float hix = S.window_clamp ? S.gmax_x + S.o[0] : S.gmax_x;
  const float hiz = S.window_clamp ? S.gmax_z + S.o[2] : S.gmax_z;
  if (np[0] < lox) {
    np[0] = lox;
    nv[0] = fmaxf(nv[0], 0.0f) * 0.5f;
  } else if (np[0] > hix) {
    np[0] = hix;
    nv[0] = fminf(nv[0], 0.0f) * 0.5f;
  }
  if (np[2] < loz) {
    np[2] = loz;
    nv[2] = fmaxf(nv[2], 0.0f) * 0.5f;
  } else if (np[2] > hiz) {
    np[2] = hiz;
    nv[2] = fminf(nv[2], 0.0f) * 0.5f;
  }
  // x, v and C are final here: store them before the F update so they are
  // not live across the snow SVD.
  at4(nxt.x[0], os) = np[0];
  at4(nxt.x[1], os) = np[1];
  at4(nxt.x[2], os) = np[2];
  if constexpr (COUNT) count_next(s, S, np[0], np[1], np[2], j, n);
  at4(nxt.v[0], os) = nv[0] * S.damping;
  at4(nxt.v[1], os) = nv[1] * S.damping;
  at4(nxt.v[2], os) = nv[2] * S.damping;
  if constexpr (STORE == 0) {
#pragma unroll
    for (int k = 0; k < 9; k++) at4(nxt.C[k], os) = nc[k];
  } else {  // words 0-3; C[8] shares word 8 with F[8], written below
#pragma unroll
    for (int w = 0; w < 4; w++) at4(nxt.CF[w], os) = pack16(nc[2 * w], nc[2 * w + 1], STORE);
  }
  // Q9: F' = (I + dt grad v) F; validity; J = clamp(det F', 0.1, 10).
  float nf[9];
#pragma unroll
  for (int r = 0; r < 3; r++)
#pragma unroll
    for (int cc = 0; cc < 3; cc++) {
      float sum = 0.0f;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        const float kron = (r == k) ? 1.0f : 0.0f;
        const float term = kron + gv[3 * r + k] * dt;
        sum += term * F[3 * k + cc];
      }
      nf[3 * r + cc] = sum;
    }
  bool f_valid = true;
#pragma unroll
  for (int k = 0; k < 9; k++)
    if (!isfinite(nf[k]) || fabsf(nf[k]) > 1e3f) f_valid = false;
  float jp = pjp;
  float fj;
  if (f_valid) {
    const MatDev& M = mats[mi < (uint32_t)nmat ? mi : (uint32_t)nmat];
#if AEMH_DET
    if (M.model == kModelSnow) snow_clamp(M, nf, &jp, s.counters + kCntSvd64);
#else
    if (M.model == kModelSnow) snow_clamp_fast(M, nf, &jp, s.counters + kCntSvd64);
#endif
    fj = det3(nf);
    if (fj < 0.1f) fj = 0.1f;
    if (fj > 10.0f) fj = 10.0f;
  } else {
    identity3(nf);
    fj = 1.0f;
  }
  if constexpr (STORE == 0) {
#pragma unroll
    for (int k = 0; k < 9; k++) at4(nxt.F[k], os) = nf[k];
  } else {
#pragma unroll
    for (int w = 0; w < 4; w++)
      at4(nxt.CF[4 + w], os) = pack16(nf[2 * w] - eye3(2 * w), nf[2 * w + 1] - eye3(2 * w + 1), STORE);
    at4(nxt.CF[8], os) = pack16(nc[8], nf[8] - 1.0f, STORE);
  }
  at4(nxt.J, os) = fj;
  at4(nxt.Jp, os) = jp;
}

#if AEMH_DET
// ---- Reference-order FP32 P2G (the serial substep, mpm_solver.rs:831-903) --
// The reference adds each node's contributions in particle-list order
// starting from 0.0, so its float sums have exactly one order.  k_ser_emit
// computes every contribution as the reference does (same expressions,
// same skips) and keys it by (node, list position); after a radix sort of
// the keys, k_ser_sum adds each node's entries in that order.
template <int STORE>
__global__ __launch_bounds__(256) void k_ser_emit(PBuf P, const SceneDev* __restrict__ sc, int ns,
                                                  const MatDev* __restrict__ mats, int nmat,
                                                  uint32_t n, SerialDev sd, uint32_t nodes) {
  const uint32_t j = blockIdx.x * 256u + threadIdx.x;
  if (j >= n) return;
  const SceneDev S = sc[scene_of(sc, ns, j)];
  const Stencil st = make_stencil(S, P.x[0][j], P.x[1][j], P.x[2][j]);
  P2GPart q;
  p2g_inputs<STORE>(P, j, S, st, mats, nmat, !S.vel_only, q, TableLn{sd.lntab});
  const unsigned long long id = P.id[j];
  const size_t e0 = (size_t)j * 27u;
  const unsigned long long unused = (unsigned long long)nodes << sd.id_bits;
  for (int k = 0; k < 27; k++) {
    sd.key[0][e0 + k] = unused;
    sd.val[0][e0 + k] = (uint32_t)(e0 + k);
  }
  const uint32_t boff = S.block_off;
  const int nb0 = S.nb[0], nb1 = S.nb[1];
  auto put = [&](int f, int di, int dj, int dk, int gx, int gy, int gz, float v) {
    const size_t e = e0 + (size_t)((dk * 3 + dj) * 3 + di);
    if (f == 0) {
      const uint32_t b = boff + (uint32_t)(((gz >> 2) * nb1 + (gy >> 2)) * nb0 + (gx >> 2));
      const uint32_t node = (b << 6) | (uint32_t)(((gz & 3) << 4) | ((gy & 3) << 2) | (gx & 3));
      sd.key[0][e] = ((unsigned long long)node << sd.id_bits) | id;
      if (S.vel_only)  // the velocity-only tier leaves the force at zero
        for (int g = 4; g < 8; g++) sd.contrib[e * 8 + g] = 0.0f;
    }
    sd.contrib[e * 8 + f] = v;
  };
  if (S.vel_only)
    p2g_nodes<false>(S.n[0], S.n[1], S.n[2], S.dx, S.o[0], S.o[1], S.o[2], st, q, put);
  else
    p2g_nodes<true>(S.n[0], S.n[1], S.n[2], S.dx, S.o[0], S.o[1], S.o[2], st, q, put);
}

// One thread per sorted entry; the first entry of each node adds up the
// node's run in list order and writes its 7 sums.  Nodes no particle
// reaches keep the zeros the scatter kernel wrote.
__global__ __launch_bounds__(256) void k_ser_sum(SerialDev sd, uint32_t m, uint32_t nodes, GridDev G) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= m) return;
  const unsigned long long* key = sd.key[1];
  const unsigned long long node = key[i] >> sd.id_bits;
  if (node >= nodes) return;
  if (i > 0 && (key[i - 1] >> sd.id_bits) == node) return;
  float sum[7] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  for (uint32_t k = i; k < m && (key[k] >> sd.id_bits) == node; k++) {
    const float* c = sd.contrib + (size_t)sd.val[1][k] * 8;
#pragma unroll
    for (int f = 0; f < 7; f++) sum[f] += c[f];
  }
#pragma unroll
  for (int f = 0; f < 7; f++) G.accf[(size_t)f * G.nodes + node] = sum[f];
}
#endif

}  // namespace

#if AEMH_DET
void launch_p2g_serial(hipStream_t st, PBuf cur, const SceneDev* sc, int ns, const MatDev* mats,
                       int nmat, uint32_t n, SerialDev sd, GridDev g, int storage) {
  if (!n) return;
  const unsigned grid = (n + 255u) / 256u;
  const uint32_t m = n * 27u;
  if (storage == 1)
    launch_k(k_ser_emit<1>, grid, 256, 0, st, cur, sc, ns, mats, nmat, n, sd, g.nodes);
  else if (storage == 2)
    launch_k(k_ser_emit<2>, grid, 256, 0, st, cur, sc, ns, mats, nmat, n, sd, g.nodes);
  else
    launch_k(k_ser_emit<0>, grid, 256, 0, st, cur, sc, ns, mats, nmat, n, sd, g.nodes);
  launch_sort_pairs_u64(st, sd.tmp, sd.tmp_bytes, sd.key[0], sd.key[1], sd.val[0], sd.val[1], m,
                        sd.end_bit);
  launch_k(k_ser_sum, (m + 255u) / 256u, 256, 0, st, sd, m, g.nodes, g);
}
#endif

template <int STORE, bool PAIRS>
void launch_p2g_store(hipStream_t st, int tiles, int partials, int layers, unsigned grid, PBuf cur,
                      const SceneDev* sc, int ns, const MatDev* mats, int nmat, uint32_t n,
                      SortDev s, GridDev g) {
  if (tiles <= 2)
    launch_k(k_p2g<STORE, 2, PAIRS>, grid, 256, 0, st, cur, sc, ns, mats, nmat, n, s, g, partials, layers);
  else if (tiles == 3)
    launch_k(k_p2g<STORE, 3, PAIRS>, grid, 256, 0, st, cur, sc, ns, mats, nmat, n, s, g, partials, layers);
  else if (tiles == 4)
    launch_k(k_p2g<STORE, 4, PAIRS>, grid, 256, 0, st, cur, sc, ns, mats, nmat, n, s, g, partials, layers);
  else
    launch_k(k_p2g<STORE, 5, PAIRS>, grid, 256, 0, st, cur, sc, ns, mats, nmat, n, s, g, partials, layers);
}

template <bool PAIRS>
void launch_p2g_pairs(hipStream_t st, const LaunchShape& ls, PBuf cur, const SceneDev* sc, int ns,
                      const MatDev* mats, int nmat, uint32_t n, SortDev s, GridDev g, int storage) {
  const unsigned grid = (n + 255u) / 256u;
  if (storage == 1)
    launch_p2g_store<1, PAIRS>(st, ls.p2g_tiles, ls.p2g_partials, ls.p2g_layers, grid, cur, sc, ns, mats, nmat, n, s, g);
  else if (storage == 2)
    launch_p2g_store<2, PAIRS>(st, ls.p2g_tiles, ls.p2g_partials, ls.p2g_layers, grid, cur, sc, ns, mats, nmat, n, s, g);
  else
    launch_p2g_store<0, PAIRS>(st, ls.p2g_tiles, ls.p2g_partials, ls.p2g_layers, grid, cur, sc, ns, mats, nmat, n, s, g);
}

void launch_p2g(hipStream_t st, const LaunchShape& ls, PBuf cur, const SceneDev* sc, int ns,
                const MatDev* mats, int nmat, uint32_t n, SortDev s, GridDev g, int storage) {
  if (!n) return;
#if !AEMH_DET
  if (ls.p2g_pairs) {
    launch_p2g_pairs<true>(st, ls, cur, sc, ns, mats, nmat, n, s, g, storage);
    return;
  }
#endif
  launch_p2g_pairs<false>(st, ls, cur, sc, ns, mats, nmat, n, s, g, storage);
}

void launch_grid_update(hipStream_t st, const LaunchShape& ls, const SceneDev* sc, int ns,
                        SortDev s, GridDev g, uint32_t n, uint32_t nb_total) {
  launch_k(k_grid_update<kTileFields>, ls.upd_wgs, 256, 0, st, sc, ns, s, g, ls.p2g_tiles,
           ls.p2g_partials, n, nb_total);
}
#if !AEMH_DET
// The reference-order mode's update: 7 planes (m, p, f), no partials.
void launch_grid_update_mpf(hipStream_t st, const LaunchShape& ls, const SceneDev* sc, int ns,
                            SortDev s, GridDev g, uint32_t n, uint32_t nb_total) {
  launch_k(k_grid_update<7>, ls.upd_wgs, 256, 0, st, sc, ns, s, g, ls.p2g_tiles, 0, n, nb_total);
}
#endif

void launch_g2p(hipStream_t st, const LaunchShape& ls, PBuf cur, PBuf nxt, const SceneDev* sc,
                int ns, const MatDev* mats, int nmat, uint32_t n, SortDev s, GridDev g,
                int storage, int inplace, int count) {
  if (!n) return;
  const unsigned wg = (unsigned)ls.g2p_wg;
  const unsigned grid = (n + wg - 1u) / wg;
  const size_t lds = (size_t)g2p_tiles_for((int)wg) * kTileNodes * sizeof(float4);
  if (count) {
    if (storage == 1)
      launch_k(k_g2p<1, true>, grid, wg, lds, st, cur, nxt, sc, ns, mats, nmat, n, s, g, inplace);
    else if (storage == 2)
      launch_k(k_g2p<2, true>, grid, wg, lds, st, cur, nxt, sc, ns, mats, nmat, n, s, g, inplace);
    else
      launch_k(k_g2p<0, true>, grid, wg, lds, st, cur, nxt, sc, ns, mats, nmat, n, s, g, inplace);
  } else {
    if (storage == 1)
      launch_k(k_g2p<1, false>, grid, wg, lds, st, cur, nxt, sc, ns, mats, nmat, n, s, g, inplace);
    else if (storage == 2)
      launch_k(k_g2p<2, false>, grid, wg, lds, st, cur, nxt, sc, ns, mats, nmat, n, s, g, inplace);
    else
      launch_k(k_g2p<0, false>, grid, wg, lds, st, cur, nxt, sc, ns, mats, nmat, n, s, g, inplace);
  }
}

}  // namespace AEMH_NS
}  // namespace aemh
