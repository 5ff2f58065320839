// kernels.inc.h -- P2G, grid update and G2P of one MLS-MPM substep.
//
// Compiled twice: k_fast.hip defines AEMH_NS=fast, AEMH_DET=0 (FP32 atomics,
// -ffp-contract=fast); k_det.hip defines AEMH_NS=det, AEMH_DET=1 (Q32.32 int64
// atomics, -ffp-contract=off, reference evaluation order).
//
// Reference: ae_particle/src/mpm_solver.rs
//   P2G          :831-903 (serial) / :1108-1197 (rayon fold/reduce)
//   grid update  :112-139 (+ ae_compute/src/mpm_compute.rs:343-389 margin)
//   G2P          :914-1053 (serial) / :1208-1325 (parallel)
//   velocity-only:1353-1572
#include "aemh_launch.h"
#include "aemh_physics.h"

namespace aemh {
namespace AEMH_NS {
namespace {

#if AEMH_DET
using Acc = long long;
#else
using Acc = float;
#endif

// Up to this many distinct base blocks of one workgroup's particles are
// staged in LDS; the rare rest scatter straight to HBM.
constexpr int kMaxTiles = 4;
constexpr int kTileVals = 7 * kTileNodes;

__device__ __forceinline__ void acc_global(GridDev& G, int f, uint32_t node, Acc v) {
#if AEMH_DET
  atomicAdd((unsigned long long*)&G.accq[(size_t)f * G.nodes + node], (unsigned long long)v);
#else
  unsafeAtomicAdd(&G.accf[(size_t)f * G.nodes + node], v);
#endif
}
__device__ __forceinline__ void acc_lds(Acc* t, Acc v) {
#if AEMH_DET
  atomicAdd((unsigned long long*)t, (unsigned long long)v);
#else
  atomicAdd(t, v);
#endif
}

template <int STORE>
__global__ __launch_bounds__(256) void k_p2g(PBuf P, const SceneDev* __restrict__ sc, int ns,
                                             const MatDev* __restrict__ mats, int nmat,
                                             uint32_t n, SortDev s, GridDev G) {
  __shared__ Acc tile[kMaxTiles * kTileVals];
  __shared__ uint32_t skey[256];
  __shared__ int tsc[kMaxTiles];
  __shared__ int torg[kMaxTiles][3];
  __shared__ uint32_t wcnt[4];

  const int t = threadIdx.x;
  const int lane = __lane_id(), wave = t >> 6;
  const uint32_t j = blockIdx.x * 256u + (uint32_t)t;
  const bool valid = j < n;
  uint32_t p = 0, key = 0xffffffffu;
  if (valid) {
    p = s.perm[j];
    key = s.key[p];
  }
  skey[t] = key;
  __syncthreads();
  const bool head = valid && (t == 0 || skey[t - 1] != key);
  const unsigned long long hb = __ballot(head);
  const unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  if (lane == 0) wcnt[wave] = (uint32_t)__popcll(hb);
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < 4; w++) {
    if (w < wave) pre += wcnt[w];
    tot += wcnt[w];
  }
  const int slot = (int)(pre + (uint32_t)__popcll(hb & le)) - 1;
  const int ntiles = (int)min(tot, (uint32_t)kMaxTiles);

  int sid = 0;
  Stencil st = {{0, 0, 0}, {0.f, 0.f, 0.f}};
  if (valid) {
    sid = scene_of(sc, ns, j);
    st = make_stencil(sc[sid], P.x[0][p], P.x[1][p], P.x[2][p]);
  }
  if (head && slot < kMaxTiles) {
    const SceneDev& S = sc[sid];
    tsc[slot] = sid;
#pragma unroll
    for (int a = 0; a < 3; a++) torg[slot][a] = (min(st.g[a], S.n[a] - 1) >> 2) * 4;
  }
  for (int e = t; e < ntiles * kTileVals; e += 256) tile[e] = 0;
  __syncthreads();

  [[maybe_unused]] unsigned long long qmax = 0;
  if (valid) {
    const SceneDev& S = sc[sid];
    const bool use_tile = slot < kMaxTiles;
    int og[3] = {0, 0, 0};
    if (use_tile) {
      og[0] = torg[slot][0];
      og[1] = torg[slot][1];
      og[2] = torg[slot][2];
    }
    Acc* tl = tile + (use_tile ? slot : 0) * kTileVals;
    const float px = P.x[0][p], py = P.x[1][p], pz = P.x[2][p];
    const float vx = P.v[0][p], vy = P.v[1][p], vz = P.v[2][p];
    const float mass = P.mass[p];
    float wx[3], wy[3], wz[3], dwx[3], dwy[3], dwz[3];
#pragma unroll
    for (int a = 0; a < 3; a++) {
      wx[a] = b_spline((float)a - st.fx[0]);
      wy[a] = b_spline((float)a - st.fx[1]);
      wz[a] = b_spline((float)a - st.fx[2]);
      dwx[a] = b_spline_deriv((float)a - st.fx[0]) * S.inv_dx;
      dwy[a] = b_spline_deriv((float)a - st.fx[1]) * S.inv_dx;
      dwz[a] = b_spline_deriv((float)a - st.fx[2]) * S.inv_dx;
    }
    float c[9], sv[9];
    const bool full = !S.vel_only;
    if (full) {
      float F[9];
#pragma unroll
      for (int k = 0; k < 9; k++) {
        c[k] = get_c(P, p, k, STORE);
        F[k] = get_f(P, p, k, STORE);
      }
      const uint32_t mi = P.mat[p];
      const MatDev& M = mats[mi < (uint32_t)nmat ? mi : (uint32_t)nmat];
      float stress[9];
      compute_stress(M, F, P.J[p], P.Jp[p], stress);
      const float vol = mass / M.density;
#pragma unroll
      for (int k = 0; k < 9; k++) sv[k] = vol * stress[k];
    }
    auto put = [&](int f, int lx, int ly, int lz, int gx, int gy, int gz, float val) {
#if AEMH_DET
      const Acc a = fp_from_f32(val);
      const unsigned long long mag = a < 0 ? 0ull - (unsigned long long)a : (unsigned long long)a;
      qmax = mag > qmax ? mag : qmax;
      if (a == 0) return;
#else
      const Acc a = val;
#endif
      if (use_tile && (unsigned)lx < 6u && (unsigned)ly < 6u && (unsigned)lz < 6u)
        acc_lds(&tl[f * kTileNodes + (lz * 6 + ly) * 6 + lx], a);
      else
        acc_global(G, f, node_addr(S, gx, gy, gz), a);
    };
#pragma unroll
    for (int dk = 0; dk < 3; dk++)
#pragma unroll
      for (int dj = 0; dj < 3; dj++)
#pragma unroll
        for (int di = 0; di < 3; di++) {
          // velocity-only tier multiplies wx * (wy * wz) (mpm_solver.rs:1425-1428)
          const float w = full ? wx[di] * wy[dj] * wz[dk] : wx[di] * (wy[dj] * wz[dk]);
          if (w < 1e-10f) continue;
          const int gx = st.g[0] + di, gy = st.g[1] + dj, gz = st.g[2] + dk;
          if (gx >= S.n[0] || gy >= S.n[1] || gz >= S.n[2]) continue;
          const int lx = gx - og[0], ly = gy - og[1], lz = gz - og[2];
          if (full) {
            const float dxg = ((float)gx + 0.5f) * S.dx + S.o[0] - px;
            const float dyg = ((float)gy + 0.5f) * S.dx + S.o[1] - py;
            const float dzg = ((float)gz + 0.5f) * S.dx + S.o[2] - pz;
            const float ux = vx + c[0] * dxg + c[1] * dyg + c[2] * dzg;
            const float uy = vy + c[3] * dxg + c[4] * dyg + c[5] * dzg;
            const float uz = vz + c[6] * dxg + c[7] * dyg + c[8] * dzg;
            const float wm = w * mass;
            put(0, lx, ly, lz, gx, gy, gz, wm);
            put(1, lx, ly, lz, gx, gy, gz, w * mass * ux);
            put(2, lx, ly, lz, gx, gy, gz, w * mass * uy);
            put(3, lx, ly, lz, gx, gy, gz, w * mass * uz);
            put(4, lx, ly, lz, gx, gy, gz, sv[0] * dwx[di] + sv[1] * dwy[dj] + sv[2] * dwz[dk]);
            put(5, lx, ly, lz, gx, gy, gz, sv[3] * dwx[di] + sv[4] * dwy[dj] + sv[5] * dwz[dk]);
            put(6, lx, ly, lz, gx, gy, gz, sv[6] * dwx[di] + sv[7] * dwy[dj] + sv[8] * dwz[dk]);
          } else {
            put(0, lx, ly, lz, gx, gy, gz, w * mass);
            put(1, lx, ly, lz, gx, gy, gz, w * mass * vx);
            put(2, lx, ly, lz, gx, gy, gz, w * mass * vy);
            put(3, lx, ly, lz, gx, gy, gz, w * mass * vz);
          }
        }
  }
#if AEMH_DET
  // Largest |contribution| so the host can prove no node sum wrapped.
  for (int d = 32; d > 0; d >>= 1) {
    unsigned long long o = __shfl_xor(qmax, d, 64);
    qmax = o > qmax ? o : qmax;
  }
  if (lane == 0 && qmax) atomicMax(s.qmax, qmax);
#endif
  __syncthreads();
  // Flush the staged halos: one global atomic per non-zero node value.
  for (int e = t; e < ntiles * kTileVals; e += 256) {
    const Acc v = tile[e];
    if (v == 0) continue;
    const int ti = e / kTileVals;
    const int r = e - ti * kTileVals;
    const int f = r / kTileNodes;
    const int nn = r - f * kTileNodes;
    const int lz = nn / 36, ly = (nn / 6) % 6, lx = nn % 6;
    const SceneDev& S = sc[tsc[ti]];
    const int gx = torg[ti][0] + lx, gy = torg[ti][1] + ly, gz = torg[ti][2] + lz;
    if (gx >= S.n[0] || gy >= S.n[1] || gz >= S.n[2]) continue;
    acc_global(G, f, node_addr(S, gx, gy, gz), v);
  }
}

// Grid update over the active blocks, one wave per 4^3 block (persistent).
__global__ __launch_bounds__(256) void k_grid_update(const SceneDev* __restrict__ sc, int ns,
                                                     SortDev s, GridDev G) {
  const uint32_t lane = __lane_id();
  const uint32_t wave = (blockIdx.x * 256u + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * 256u) >> 6;
  const uint32_t nact = s.counters[0];
  for (uint32_t a = wave; a < nact; a += nwaves) {
    const uint32_t b = s.active[a];
    const SceneDev& S = sc[scene_of_block(sc, ns, b)];
    const uint32_t lb = b - S.block_off;
    const int bx = (int)(lb % (uint32_t)S.nb[0]);
    const int by = (int)((lb / (uint32_t)S.nb[0]) % (uint32_t)S.nb[1]);
    const int bz = (int)(lb / (uint32_t)(S.nb[0] * S.nb[1]));
    const int gx = bx * 4 + (int)(lane & 3), gy = by * 4 + (int)((lane >> 2) & 3),
              gz = bz * 4 + (int)(lane >> 4);
    const uint32_t node = (b << 6) | lane;
    float acc[7];
#pragma unroll
    for (int f = 0; f < 7; f++) {
#if AEMH_DET
      acc[f] = fp_to_f32(G.accq[(size_t)f * G.nodes + node]);
#else
      acc[f] = G.accf[(size_t)f * G.nodes + node];
#endif
    }
    const float m = acc[0];
    float v0 = acc[1], v1 = acc[2], v2 = acc[3];
    if (m > 1e-12f) {
      v0 /= m;
      v1 /= m;
      v2 /= m;
      v0 += (S.g[0] + acc[4] / m) * S.dt;
      v1 += (S.g[1] + acc[5] / m) * S.dt;
      v2 += (S.g[2] + acc[6] / m) * S.dt;
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

// G2P: sorted position j reads working slot perm[j] of `cur` and writes slot j
// of `nxt`, so the particle arrays come out ordered by base block.
template <int STORE>
__global__ __launch_bounds__(256) void k_g2p(PBuf cur, PBuf nxt, const SceneDev* __restrict__ sc,
                                             int ns, const MatDev* __restrict__ mats, int nmat,
                                             uint32_t n, SortDev s, GridDev G) {
  const uint32_t j = blockIdx.x * 256u + threadIdx.x;
  if (j >= n) return;
  const uint32_t p = s.perm[j];
  const SceneDev& S = sc[scene_of(sc, ns, j)];
  const float px = cur.x[0][p], py = cur.x[1][p], pz = cur.x[2][p];
  const Stencil st = make_stencil(S, px, py, pz);
  float wx[3], wy[3], wz[3];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    wx[a] = b_spline((float)a - st.fx[0]);
    wy[a] = b_spline((float)a - st.fx[1]);
    wz[a] = b_spline((float)a - st.fx[2]);
  }
  const float dt = S.dt;
  const uint16_t mi = cur.mat[p];
  nxt.mass[j] = cur.mass[p];
  nxt.mat[j] = mi;
  nxt.id[j] = cur.id[p];

  if (S.vel_only) {  // mpm_solver.rs:1465-1571
    float nv0 = 0.0f, nv1 = 0.0f, nv2 = 0.0f;
#pragma unroll
    for (int dk = 0; dk < 3; dk++)
#pragma unroll
      for (int dj = 0; dj < 3; dj++)
#pragma unroll
        for (int di = 0; di < 3; di++) {
          const float w = wx[di] * (wy[dj] * wz[dk]);
          if (w < 1e-10f) continue;
          const int gx = st.g[0] + di, gy = st.g[1] + dj, gz = st.g[2] + dk;
          if (gx >= S.n[0] || gy >= S.n[1] || gz >= S.n[2]) continue;
          const float4 g = G.gv[node_addr(S, gx, gy, gz)];
          if (g.w < 1e-12f) continue;
          nv0 += w * g.x;
          nv1 += w * g.y;
          nv2 += w * g.z;
        }
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
    nxt.x[0][j] = x0;
    nxt.x[1][j] = x1;
    nxt.x[2][j] = x2;
    nxt.v[0][j] = nv0 * S.damping;
    nxt.v[1][j] = nv1 * S.damping;
    nxt.v[2][j] = nv2 * S.damping;
#pragma unroll
    for (int k = 0; k < 9; k++) {
      put_c(nxt, j, k, get_c(cur, p, k, STORE), STORE);
      put_f(nxt, j, k, get_f(cur, p, k, STORE), STORE);
    }
    nxt.J[j] = cur.J[p];
    nxt.Jp[j] = cur.Jp[p];
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
  float nv[3] = {0.0f, 0.0f, 0.0f};
  float nc[9] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  float gv[9] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int dk = 0; dk < 3; dk++)
#pragma unroll
    for (int dj = 0; dj < 3; dj++)
#pragma unroll
      for (int di = 0; di < 3; di++) {
        const float w = wx[di] * wy[dj] * wz[dk];
        if (w < 1e-10f) continue;
        const int gx = st.g[0] + di, gy = st.g[1] + dj, gz = st.g[2] + dk;
        if (gx >= S.n[0] || gy >= S.n[1] || gz >= S.n[2]) continue;
        const float4 g = G.gv[node_addr(S, gx, gy, gz)];
        if (g.w < 1e-12f) continue;
        nv[0] += w * g.x;
        nv[1] += w * g.y;
        nv[2] += w * g.z;
        const float scale = S.four_inv_dx2 * w;
        const float d[3] = {dxa[di], dya[dj], dza[dk]};
        const float dw[3] = {dwx[di], dwy[dj], dwz[dk]};
        const float gg[3] = {g.x, g.y, g.z};
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int cc = 0; cc < 3; cc++) nc[3 * r + cc] += scale * gg[r] * d[cc];
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
          for (int cc = 0; cc < 3; cc++) gv[3 * r + cc] -= w * gg[r] * dw[cc];
      }
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
  // Q9: F' = (I + dt grad v) F; validity; J = clamp(det F', 0.1, 10).
  float F[9], nf[9];
#pragma unroll
  for (int k = 0; k < 9; k++) F[k] = get_f(cur, p, k, STORE);
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
  float jp = cur.Jp[p];
  float fj;
  if (f_valid) {
    const MatDev& M = mats[mi < (uint32_t)nmat ? mi : (uint32_t)nmat];
    if (M.model == kModelSnow) snow_clamp(M, nf, &jp);
    fj = det3(nf);
    if (fj < 0.1f) fj = 0.1f;
    if (fj > 10.0f) fj = 10.0f;
  } else {
    identity3(nf);
    fj = 1.0f;
  }
  nxt.x[0][j] = np[0];
  nxt.x[1][j] = np[1];
  nxt.x[2][j] = np[2];
  nxt.v[0][j] = nv[0] * S.damping;
  nxt.v[1][j] = nv[1] * S.damping;
  nxt.v[2][j] = nv[2] * S.damping;
#pragma unroll
  for (int k = 0; k < 9; k++) {
    put_c(nxt, j, k, nc[k], STORE);
    put_f(nxt, j, k, nf[k], STORE);
  }
  nxt.J[j] = fj;
  nxt.Jp[j] = jp;
}

}  // namespace

void launch_p2g(hipStream_t st, const LaunchShape&, PBuf cur, const SceneDev* sc, int ns,
                const MatDev* mats, int nmat, uint32_t n, SortDev s, GridDev g, int storage) {
  if (!n) return;
  const unsigned grid = (n + 255u) / 256u;
  if (storage == 1)
    k_p2g<1><<<grid, 256, 0, st>>>(cur, sc, ns, mats, nmat, n, s, g);
  else if (storage == 2)
    k_p2g<2><<<grid, 256, 0, st>>>(cur, sc, ns, mats, nmat, n, s, g);
  else
    k_p2g<0><<<grid, 256, 0, st>>>(cur, sc, ns, mats, nmat, n, s, g);
}

void launch_grid_update(hipStream_t st, const LaunchShape& ls, const SceneDev* sc, int ns,
                        SortDev s, GridDev g) {
  k_grid_update<<<ls.upd_wgs, 256, 0, st>>>(sc, ns, s, g);
}

void launch_g2p(hipStream_t st, const LaunchShape&, PBuf cur, PBuf nxt, const SceneDev* sc,
                int ns, const MatDev* mats, int nmat, uint32_t n, SortDev s, GridDev g,
                int storage) {
  if (!n) return;
  const unsigned grid = (n + 255u) / 256u;
  if (storage == 1)
    k_g2p<1><<<grid, 256, 0, st>>>(cur, nxt, sc, ns, mats, nmat, n, s, g);
  else if (storage == 2)
    k_g2p<2><<<grid, 256, 0, st>>>(cur, nxt, sc, ns, mats, nmat, n, s, g);
  else
    k_g2p<0><<<grid, 256, 0, st>>>(cur, nxt, sc, ns, mats, nmat, n, s, g);
}

}  // namespace AEMH_NS
}  // namespace aemh
