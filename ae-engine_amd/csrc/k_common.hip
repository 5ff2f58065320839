// k_common.hip -- kernels shared by both arithmetic modes: counting sort of
// particles by base block, active-block bookkeeping, grid clears, layout
// conversions between the reference's AoS arrays and HBM structure of
// arrays, parity readback and FP64 reductions.
#include "aemh_dev.h"
#include "aemh_launch.h"

namespace aemh {
namespace {

// Key + rank: every working slot computes the global block of its clamped
// base cell; runs of equal keys inside a wave take one atomic on the block
// counter, so nearly sorted input costs about one atomic per wave.
__global__ __launch_bounds__(256) void k_count(PBuf cur, const SceneDev* __restrict__ sc, int ns,
                                               uint32_t n, SortDev s) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  const bool valid = i < n;
  if (__ballot(valid) == 0) return;
  const int lane = __lane_id();
  uint32_t key = 0xffffffffu;
  if (valid) {
    const SceneDev S = sc[scene_of(sc, ns, i)];
    const Stencil st = make_stencil(S, cur.x[0][i], cur.x[1][i], cur.x[2][i]);
    key = block_key(S, st);
    s.key[i] = key;
  }
  const uint32_t prev = __shfl_up(key, 1, 64);
  const bool head = valid && (lane == 0 || prev != key);
  const unsigned long long heads = __ballot(head);
  const int nvalid = __popcll(__ballot(valid));
  const unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  uint32_t old = 0;
  if (head) {
    const unsigned long long above = heads & ~le;
    const int next = above ? (__ffsll((long long)above) - 1) : nvalid;
    old = atomicAdd(&s.cnt[key], (uint32_t)(next - lane));
  }
  const unsigned long long mine = valid ? (heads & le) : 1ull;
  const int h = 63 - __clzll((long long)mine);
  const uint32_t base = __shfl(old, h, 64);
  if (valid) s.rank[i] = base + (uint32_t)(lane - h);
  if (lane == 0) count_runs(s, blockIdx.x * 4u + (threadIdx.x >> 6), (uint32_t)__popcll(heads));
}

// Grids with few node blocks (a coarse tier grid holding half a million
// particles, say): unsorted input would queue hundreds of lanes on each block
// counter in HBM.  Each workgroup takes 4 rounds of 256 particles, counts
// them in LDS bins (a run of equal keys inside a wave still takes one LDS
// atomic), then adds each touched bin to its block counter with one HBM
// atomic; a particle's rank is that atomic's result plus its LDS rank.
constexpr uint32_t kCountLdsBins = 8192;
constexpr int kCountRounds = 4;

__global__ __launch_bounds__(256) void k_count_lds(PBuf cur, const SceneDev* __restrict__ sc,
                                                   int ns, uint32_t n, SortDev s, uint32_t nb) {
  __shared__ uint32_t bins[kCountLdsBins];
  for (uint32_t j = threadIdx.x; j < nb; j += 256u) bins[j] = 0u;
  __syncthreads();
  const int lane = __lane_id();
  const unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  uint32_t keys[kCountRounds], loc[kCountRounds];
  bool own[kCountRounds];
  uint32_t nruns = 0;
#pragma unroll
  for (int r = 0; r < kCountRounds; r++) {
    const uint32_t i = (blockIdx.x * kCountRounds + r) * 256u + threadIdx.x;
    const bool valid = i < n;
    uint32_t key = 0xffffffffu;
    if (valid) {
      const SceneDev S = sc[scene_of(sc, ns, i)];
      const Stencil st = make_stencil(S, cur.x[0][i], cur.x[1][i], cur.x[2][i]);
      key = block_key(S, st);
      s.key[i] = key;
    }
    const uint32_t prev = __shfl_up(key, 1, 64);
    const bool head = valid && (lane == 0 || prev != key);
    const unsigned long long heads = __ballot(head);
    const int nvalid = __popcll(__ballot(valid));
    nruns += (uint32_t)__popcll(heads);
    uint32_t old = 0;
    if (head) {
      const unsigned long long above = heads & ~le;
      const int next = above ? (__ffsll((long long)above) - 1) : nvalid;
      old = atomicAdd(&bins[key], (uint32_t)(next - lane));
    }
    const unsigned long long mine = valid ? (heads & le) : 1ull;
    const int h = 63 - __clzll((long long)mine);
    const uint32_t base = __shfl(old, h, 64);
    keys[r] = key;
    loc[r] = base + (uint32_t)(lane - h);
    own[r] = head && old == 0u;  // first adder of this bin in the workgroup
  }
  if (lane == 0) count_runs(s, blockIdx.x * 4u + (threadIdx.x >> 6), nruns);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kCountRounds; r++)
    if (own[r]) bins[keys[r]] = atomicAdd(&s.cnt[keys[r]], bins[keys[r]]);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kCountRounds; r++) {
    const uint32_t i = (blockIdx.x * kCountRounds + r) * 256u + threadIdx.x;
    if (i < n) s.rank[i] = bins[keys[r]] + loc[r];
  }
}

// Grids with more node blocks than k_count_lds bins.  A nearly block-sorted
// working set (re-sorted, or in place a few steps after a re-sort) gives
// each workgroup's 1024 keys a short range, so they are counted in LDS bins
// at key - kmin and each touched bin takes one HBM atomic: a few returning
// atomics per workgroup instead of one per run of every wave, whose
// latency bounds k_count.  A workgroup whose keys span more than
// kCountWinBins blocks counts with k_count's per-wave run atomics.
constexpr uint32_t kCountWinBins = 2048;

__global__ __launch_bounds__(256) void k_count_win(PBuf cur, const SceneDev* __restrict__ sc,
                                                   int ns, uint32_t n, SortDev s) {
  __shared__ uint32_t bins[kCountWinBins];
  __shared__ uint32_t wlo[4], whi[4];
  const int lane = __lane_id(), wave = (int)(threadIdx.x >> 6);
  const unsigned long long le = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  uint32_t keys[kCountRounds];
  uint32_t lo = 0xffffffffu, hi = 0u;
#pragma unroll
  for (int r = 0; r < kCountRounds; r++) {
    const uint32_t i = (blockIdx.x * kCountRounds + r) * 256u + threadIdx.x;
    uint32_t key = 0xffffffffu;
    if (i < n) {
      const SceneDev S = sc[scene_of(sc, ns, i)];
      const Stencil st = make_stencil(S, cur.x[0][i], cur.x[1][i], cur.x[2][i]);
      key = block_key(S, st);
      s.key[i] = key;
      lo = min(lo, key);
      hi = max(hi, key);
    }
    keys[r] = key;
  }
#pragma unroll
  for (int d = 32; d > 0; d >>= 1) {
    lo = min(lo, (uint32_t)__shfl_xor((int)lo, d, 64));
    hi = max(hi, (uint32_t)__shfl_xor((int)hi, d, 64));
  }
  if (lane == 0) {
    wlo[wave] = lo;
    whi[wave] = hi;
  }
  __syncthreads();
  const uint32_t kmin = min(min(wlo[0], wlo[1]), min(wlo[2], wlo[3]));
  const uint32_t kmax = max(max(whi[0], whi[1]), max(whi[2], whi[3]));
  // The workgroup's first particle is below n, so kmin <= kmax.
  if (kmax - kmin >= kCountWinBins) {
    uint32_t nruns = 0;
#pragma unroll
    for (int r = 0; r < kCountRounds; r++) {
      const uint32_t i = (blockIdx.x * kCountRounds + r) * 256u + threadIdx.x;
      const bool valid = i < n;
      const uint32_t key = keys[r];
      const uint32_t prev = __shfl_up(key, 1, 64);
      const bool head = valid && (lane == 0 || prev != key);
      const unsigned long long heads = __ballot(head);
      const int nvalid = __popcll(__ballot(valid));
      nruns += (uint32_t)__popcll(heads);
      uint32_t old = 0;
      if (head) {
        const unsigned long long above = heads & ~le;
        const int next = above ? (__ffsll((long long)above) - 1) : nvalid;
        old = atomicAdd(&s.cnt[key], (uint32_t)(next - lane));
      }
      const unsigned long long mine = valid ? (heads & le) : 1ull;
      const int h = 63 - __clzll((long long)mine);
      const uint32_t base = __shfl(old, h, 64);
      if (valid) s.rank[i] = base + (uint32_t)(lane - h);
    }
    if (lane == 0) count_runs(s, blockIdx.x * 4u + (uint32_t)wave, nruns);
    return;  // the whole workgroup takes this branch
  }
  for (uint32_t b = threadIdx.x; b <= kmax - kmin; b += 256u) bins[b] = 0u;
  __syncthreads();
  uint32_t loc[kCountRounds];
  bool own[kCountRounds];
  uint32_t nruns = 0;
#pragma unroll
  for (int r = 0; r < kCountRounds; r++) {
    const uint32_t i = (blockIdx.x * kCountRounds + r) * 256u + threadIdx.x;
    const bool valid = i < n;
    const uint32_t key = keys[r];
    const uint32_t prev = __shfl_up(key, 1, 64);
    const bool head = valid && (lane == 0 || prev != key);
    const unsigned long long heads = __ballot(head);
    const int nvalid = __popcll(__ballot(valid));
    nruns += (uint32_t)__popcll(heads);
    uint32_t old = 0;
    if (head) {
      const unsigned long long above = heads & ~le;
      const int next = above ? (__ffsll((long long)above) - 1) : nvalid;
      old = atomicAdd(&bins[key - kmin], (uint32_t)(next - lane));
    }
    const unsigned long long mine = valid ? (heads & le) : 1ull;
    const int h = 63 - __clzll((long long)mine);
    const uint32_t base = __shfl(old, h, 64);
    loc[r] = base + (uint32_t)(lane - h);
    own[r] = head && old == 0u;  // first adder of this bin in the workgroup
  }
  if (lane == 0) count_runs(s, blockIdx.x * 4u + (uint32_t)wave, nruns);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kCountRounds; r++)
    if (own[r]) bins[keys[r] - kmin] = atomicAdd(&s.cnt[keys[r]], bins[keys[r] - kmin]);
  __syncthreads();
#pragma unroll
  for (int r = 0; r < kCountRounds; r++) {
    const uint32_t i = (blockIdx.x * kCountRounds + r) * 256u + threadIdx.x;
    if (i < n) s.rank[i] = bins[keys[r] - kmin] + loc[r];
  }
}

// Every non-empty base block activates the 2x2x2 node blocks its 6^3 halo
// can touch (idempotent plain stores, no atomics).
__device__ __forceinline__ void activate_halo(const SceneDev* __restrict__ sc, int ns, SortDev& s,
                                              uint32_t b) {
  const SceneDev S = sc[scene_of_block(sc, ns, b)];
  const uint32_t lb = b - S.block_off;
  const int bx = (int)(lb % (uint32_t)S.nb[0]);
  const int by = (int)((lb / (uint32_t)S.nb[0]) % (uint32_t)S.nb[1]);
  const int bz = (int)(lb / (uint32_t)(S.nb[0] * S.nb[1]));
  for (int dz = 0; dz < 2; dz++)
    for (int dy = 0; dy < 2; dy++)
      for (int dx = 0; dx < 2; dx++) {
        const int x = bx + dx, y = by + dy, z = bz + dz;
        if (x >= S.nb[0] || y >= S.nb[1] || z >= S.nb[2]) continue;
        s.flag[S.block_off + (uint32_t)((z * S.nb[1] + y) * S.nb[0] + x)] = 1u;
      }
}

// The substep's bookkeeping, before the scatter counts the active blocks:
// the per-substep counters (not the sticky overflow flag) are zeroed, and the
// count's wave runs go to the host-mapped word (their counters cleared for
// the next count).  One workgroup.
__device__ __forceinline__ void start_substep(SortDev& s) {
  if (threadIdx.x < kCntOverflow) s.counters[threadIdx.x] = 0u;
  if (threadIdx.x < 64) {
    uint32_t r = s.runs[threadIdx.x * kRunStride];
    s.runs[threadIdx.x * kRunStride] = 0u;
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) r += __shfl_down(r, d, 64);
    if (threadIdx.x == 0 && s.runs_host) *(volatile uint32_t*)s.runs_host = r;
  }
}

// At most this many scan tiles (65,536 blocks): the scatter's waves scan the
// tile sums themselves, and the count scan's first workgroup does the
// substep's bookkeeping, so no tile-sum kernel runs.
constexpr uint32_t kScatterScanTiles = 64;

// Exclusive scan of the per-block counts, tile of 1024 per workgroup; zeroes
// the counts for the next substep and activates the halos of non-empty blocks.
__global__ __launch_bounds__(256) void k_scan_tiles(const SceneDev* __restrict__ sc, int ns,
                                                    SortDev s, uint32_t nb, int book) {
  if (book && blockIdx.x == 0) start_substep(s);
  __shared__ uint32_t wtot[4];
  const uint32_t base = blockIdx.x * kScanTile + threadIdx.x * 4u;
  uint32_t v[4];
  for (int k = 0; k < 4; k++) {
    v[k] = (base + k < nb) ? s.cnt[base + k] : 0u;
    if (base + k < nb) s.cnt[base + k] = 0u;
    if (v[k]) activate_halo(sc, ns, s, base + k);
  }
  const uint32_t sum = v[0] + v[1] + v[2] + v[3];
  const uint32_t incl = wave_incl_scan(sum);
  const int wave = threadIdx.x >> 6, lane = __lane_id();
  if (lane == 63) wtot[wave] = incl;
  __syncthreads();
  uint32_t wpre = 0;
  for (int w = 0; w < wave; w++) wpre += wtot[w];
  uint32_t run = wpre + incl - sum;
  for (int k = 0; k < 4; k++) {
    if (base + k < nb) s.off[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 255) s.traw[blockIdx.x] = wpre + incl;
}

// Exclusive scan of the tile sums (one workgroup), for more than
// kScatterScanTiles tiles, with the substep's bookkeeping.
__global__ __launch_bounds__(256) void k_scan_tsum(SortDev s, uint32_t ntiles) {
  __shared__ uint32_t wtot[4];
  __shared__ uint32_t carry_s;
  if (threadIdx.x == 0) carry_s = 0;
  start_substep(s);
  __syncthreads();
  for (uint32_t c = 0; c < ntiles; c += 256) {
    const uint32_t i = c + threadIdx.x;
    const uint32_t v = i < ntiles ? s.traw[i] : 0u;
    const uint32_t incl = wave_incl_scan(v);
    const int wave = threadIdx.x >> 6, lane = __lane_id();
    if (lane == 63) wtot[wave] = incl;
    __syncthreads();
    uint32_t wpre = 0, tot = 0;
    for (int w = 0; w < 4; w++) {
      if (w < wave) wpre += wtot[w];
      tot += wtot[w];
    }
    const uint32_t carry = carry_s;
    if (i < ntiles) s.tsum[i] = carry + wpre + incl - v;
    __syncthreads();
    if (threadIdx.x == 0) carry_s = carry + tot;
    __syncthreads();
  }
}

// Particle slots into block order (perm), fused with the compaction of the
// activation flags into the active-block list (one atomic per wave; flags
// cleared for the next substep) and the zeroing of each listed block's
// accumulators, which the wave does one block at a time with coalesced
// stores.  All of it only needs the finished scans.  With at most
// kScatterScanTiles tiles every wave scans the tile sums itself (and the
// first wave stores the scan for the later kernels).
__global__ __launch_bounds__(256) void k_scatter(uint32_t n, uint32_t nb, SortDev s, GridDev g,
                                                 int acc, uint32_t ntiles) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  const bool on = i < nb && s.flag[i] != 0u;
  const unsigned long long m = __ballot(on);
  if (m) {  // wave-uniform
    const int lane = __lane_id();
    const int leader = __ffsll((long long)m) - 1;
    uint32_t base = 0;
    if (lane == leader) base = atomicAdd(&s.counters[kCntActive], (uint32_t)__popcll(m));
    base = __shfl(base, leader, 64);
    if (on) {
      const unsigned long long lt = (1ull << lane) - 1ull;
      s.active[base + (uint32_t)__popcll(m & lt)] = i;
      s.flag[i] = 0u;
    }
    const uint32_t b0 = i - (uint32_t)lane;  // block of lane 0
    for (unsigned long long r = m; r; r &= r - 1ull) {
      const uint32_t node = ((b0 + (uint32_t)(__ffsll((long long)r) - 1)) << 6) | (uint32_t)lane;
      if (acc == kAccQ32) {
        for (int f = 0; f < kDetFields; f++) g.accq[(size_t)f * g.nodes + node] = 0;
      } else {
        const int planes = acc == kAccF32Fold ? 4 : 7;
        for (int f = 0; f < planes; f++) g.accf[(size_t)f * g.nodes + node] = 0.0f;
      }
    }
  }
  // key, rank and block offset first, so the tile prefix below does not
  // add a round trip to memory
  uint32_t k = 0, r = 0, o = 0;
  if (i < n) {
    k = s.key[i];
    r = s.rank[i];
    o = s.off[k];
  }
  uint32_t tpre;
  if (ntiles <= kScatterScanTiles) {  // uniform; every lane active here
    const int lane = __lane_id();
    const uint32_t v = lane < (int)ntiles ? s.traw[lane] : 0u;
    const uint32_t excl = wave_incl_scan(v) - v;
    if (blockIdx.x == 0 && threadIdx.x < ntiles) s.tsum[threadIdx.x] = excl;
    tpre = (uint32_t)__shfl((int)excl, (int)(k / kScanTile), 64);
  } else {
    tpre = i < n ? s.tsum[k / kScanTile] : 0u;
  }
  if (i < n) s.perm[o + tpre + r] = i;
}

// Zero the accumulators of this substep's active blocks (one wave per block).
__global__ __launch_bounds__(256) void k_clear(SortDev s, GridDev g, int acc) {
  const uint32_t lane = __lane_id();
  const uint32_t wave = (blockIdx.x * 256u + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * 256u) >> 6;
  const uint32_t nact = s.counters[0];
  for (uint32_t a = wave; a < nact; a += nwaves) {
    const uint32_t b = s.active[a];
    const uint32_t node = (b << 6) | lane;
    if (acc == kAccQ32) {
      for (int f = 0; f < kDetFields; f++) g.accq[(size_t)f * g.nodes + node] = 0;
    } else {
      const int planes = acc == kAccF32Fold ? 4 : 7;
      for (int f = 0; f < planes; f++) g.accf[(size_t)f * g.nodes + node] = 0.0f;
    }
  }
}

// Working slot j <- canonical particle idx[j] (canonical arrays are FP32).
__global__ __launch_bounds__(256) void k_gather_work(PBuf canon, PBuf work, const uint32_t* idx,
                                                     uint32_t n, int storage) {
  const uint32_t j = blockIdx.x * 256u + threadIdx.x;
  if (j >= n) return;
  const uint32_t p = idx[j];
  for (int a = 0; a < 3; a++) {
    work.x[a][j] = canon.x[a][p];
    work.v[a][j] = canon.v[a][p];
  }
  float c[9], f[9];
  for (int k = 0; k < 9; k++) {
    c[k] = canon.C[k][p];
    f[k] = canon.F[k][p];
  }
  put_c9(work, j, c, storage);
  put_f9(work, j, f, c[8], storage);
  work.J[j] = canon.J[p];
  work.Jp[j] = canon.Jp[p];
  work.mass[j] = canon.mass[p];
  work.mat[j] = canon.mat[p];
  work.id[j] = j;
}

__global__ __launch_bounds__(256) void k_scatter_work(PBuf work, PBuf canon, const uint32_t* idx,
                                                      uint32_t n, int storage) {
  const uint32_t j = blockIdx.x * 256u + threadIdx.x;
  if (j >= n) return;
  const uint32_t p = idx[work.id[j]];
  for (int a = 0; a < 3; a++) {
    canon.x[a][p] = work.x[a][j];
    canon.v[a][p] = work.v[a][j];
  }
  float c[9], f[9];
  load_cf(work, j, c, f, storage);
  for (int k = 0; k < 9; k++) {
    canon.C[k][p] = c[k];
    canon.F[k][p] = f[k];
  }
  canon.J[p] = work.J[j];
  canon.Jp[p] = work.Jp[j];
}

__global__ __launch_bounds__(256) void k_fill(float* p, float v, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_grid_readback(const SceneDev* __restrict__ scd,
                                                       int scene, SortDev s, GridDev g, int acc,
                                                       float* m, float* p3, float* f3, float* v3,
                                                       long long* q7) {
  const SceneDev S = scd[scene];
  const uint32_t lane = __lane_id();
  const uint32_t wave = (blockIdx.x * 256u + threadIdx.x) >> 6;
  const uint32_t nwaves = (gridDim.x * 256u) >> 6;
  const uint32_t nact = s.counters[0];
  const uint32_t nscene_blocks = (uint32_t)(S.nb[0] * S.nb[1] * S.nb[2]);
  for (uint32_t a = wave; a < nact; a += nwaves) {
    const uint32_t b = s.active[a];
    if (b < S.block_off || b >= S.block_off + nscene_blocks) continue;
    const uint32_t lb = b - S.block_off;
    const int bx = (int)(lb % (uint32_t)S.nb[0]);
    const int by = (int)((lb / (uint32_t)S.nb[0]) % (uint32_t)S.nb[1]);
    const int bz = (int)(lb / (uint32_t)(S.nb[0] * S.nb[1]));
    const int gx = bx * 4 + (int)(lane & 3), gy = by * 4 + (int)((lane >> 2) & 3),
              gz = bz * 4 + (int)(lane >> 4);
    if (gx >= S.n[0] || gy >= S.n[1] || gz >= S.n[2]) continue;
    const size_t d = (size_t)gx + (size_t)gy * S.n[0] + (size_t)gz * S.n[0] * S.n[1];
    const uint32_t node = (b << 6) | lane;
    const float4 gv = g.gv[node];
    if (acc == kAccQ32) {
      float a[7];
      for (int f = 0; f < 7; f++) {
        long long r = g.accq[(size_t)f * g.nodes + node];
        if (q7) q7[7 * d + f] = r;
        a[f] = fp_to_f32(r);
      }
      if (m) m[d] = a[0];
      if (p3) for (int k = 0; k < 3; k++) p3[3 * d + k] = a[1 + k];
      if (f3) for (int k = 0; k < 3; k++) f3[3 * d + k] = a[4 + k];
    } else {
      // FP32 mode: the momentum plane holds p + f dt and there is no force plane
      const bool fold = acc == kAccF32Fold;
      if (m) m[d] = g.accf[node];
      if (p3) for (int k = 0; k < 3; k++) p3[3 * d + k] = g.accf[(size_t)(1 + k) * g.nodes + node];
      if (f3)
        for (int k = 0; k < 3; k++)
          f3[3 * d + k] = fold ? 0.0f : g.accf[(size_t)(4 + k) * g.nodes + node];
    }
    if (v3) {
      v3[3 * d + 0] = gv.x;
      v3[3 * d + 1] = gv.y;
      v3[3 * d + 2] = gv.z;
    }
  }
}

// FP64 partial sums per workgroup, then one workgroup adds the partials in a
// fixed order: deterministic reductions.
__global__ __launch_bounds__(256) void k_conservation_partial(PBuf cur, uint32_t n, double* part) {
  __shared__ double sh[5][256];
  double acc[5] = {0, 0, 0, 0, 0};
  for (uint32_t i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
    double m = cur.mass[i];
    double vx = cur.v[0][i], vy = cur.v[1][i], vz = cur.v[2][i];
    acc[0] += m;
    acc[1] += m * vx;
    acc[2] += m * vy;
    acc[3] += m * vz;
    acc[4] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
  }
  for (int k = 0; k < 5; k++) sh[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int k = 0; k < 5; k++) sh[k][threadIdx.x] += sh[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 5) part[blockIdx.x * 5 + threadIdx.x] = sh[threadIdx.x][0];
}
__global__ void k_conservation_final(const double* part, int nparts, double* out5) {
  if (threadIdx.x < 5) {
    double s = 0;
    for (int b = 0; b < nparts; b++) s += part[b * 5 + threadIdx.x];
    out5[threadIdx.x] = s;
  }
}

inline unsigned grid_for(uint32_t n, unsigned wg) { return (unsigned)((n + wg - 1) / wg); }

}  // namespace

thread_local StageEvents g_stage;

void launch_count(hipStream_t st, PBuf cur, const SceneDev* sc, int ns, uint32_t n, SortDev s,
                  uint32_t nb_total) {
  if (!n) return;
  if (nb_total <= kCountLdsBins)
    launch_k(k_count_lds, grid_for(n, 256 * kCountRounds), 256, 0, st, cur, sc, ns, n, s, nb_total);
  else
    launch_k(k_count_win, grid_for(n, 256 * kCountRounds), 256, 0, st, cur, sc, ns, n, s);
}
void launch_scan(hipStream_t st, const SceneDev* sc, int ns, SortDev s, uint32_t nb) {
  const uint32_t tiles = (nb + kScanTile - 1) / kScanTile;
  const int book = tiles <= kScatterScanTiles ? 1 : 0;
  launch_k(k_scan_tiles, tiles ? tiles : 1u, 256, 0, st, sc, ns, s, nb, book);
  if (!book) launch_k(k_scan_tsum, 1, 256, 0, st, s, tiles);
}
void launch_scatter(hipStream_t st, uint32_t n, uint32_t nb, SortDev s, GridDev g, int acc) {
  const uint32_t m = n > nb ? n : nb;
  const uint32_t tiles = (nb + kScanTile - 1) / kScanTile;
  if (m) launch_k(k_scatter, grid_for(m, 256), 256, 0, st, n, nb, s, g, acc, tiles);
}
void launch_clear(hipStream_t st, SortDev s, GridDev g, int acc, int upd_wgs) {
  k_clear<<<upd_wgs, 256, 0, st>>>(s, g, acc);
}
void launch_gather_work(hipStream_t st, PBuf canon, PBuf work, const uint32_t* idx, uint32_t n,
                        int storage) {
  if (n) k_gather_work<<<grid_for(n, 256), 256, 0, st>>>(canon, work, idx, n, storage);
}
void launch_scatter_work(hipStream_t st, PBuf work, PBuf canon, const uint32_t* idx, uint32_t n,
                         int storage) {
  if (n) k_scatter_work<<<grid_for(n, 256), 256, 0, st>>>(work, canon, idx, n, storage);
}
void launch_fill(hipStream_t st, float* p, float v, uint32_t n) {
  if (n) k_fill<<<grid_for(n, 256), 256, 0, st>>>(p, v, n);
}
void launch_grid_readback(hipStream_t st, const SceneDev* sc_d, int scene, SceneDev, SortDev s,
                          GridDev g, int acc, float* m, float* p3, float* f3, float* v3,
                          long long* q7, int upd_wgs) {
  k_grid_readback<<<upd_wgs, 256, 0, st>>>(sc_d, scene, s, g, acc, m, p3, f3, v3, q7);
}
void launch_conservation(hipStream_t st, PBuf cur, uint32_t n, double* out5) {
  k_conservation_partial<<<kConsWgs, 256, 0, st>>>(cur, n, out5 + 5);
  k_conservation_final<<<1, 64, 0, st>>>(out5 + 5, kConsWgs, out5);
}

}  // namespace aemh
