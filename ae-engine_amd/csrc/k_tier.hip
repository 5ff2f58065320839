// k_tier.hip -- device side of the LOD tier orchestrator
// (ae_engine/src/managers/simulation.rs:480-700): distance classification of
// the resident slots into near / mid / far lists, the ordered stream
// compaction that builds those lists, the near cull to the closest
// max_near_particles, the far tier's gravity-only update and the render
// filters over the cached lists.
//
// Lists keep the reference's push order (ascending slot).  Compaction is
// tile-ordered: each 1024-item tile counts its classes, one workgroup scans
// the tile counts, and each tile scatters its items in order with a
// workgroup prefix sum.
#include <hipcub/hipcub.hpp>

#include "aemh_dev.h"
#include "aemh_launch.h"

namespace aemh {
namespace {

constexpr int kSelThreads = 256;
constexpr int kSelPer = 4;  // consecutive items per thread
constexpr uint32_t kSelTile = kSelThreads * kSelPer;
constexpr int kClasses = 3;

uint32_t grid_of(uint32_t n) { return (n + 255u) / 256u; }

// Class of each active slot below count (simulation.rs:500-517): squared
// distance to the player in f32, summed x, y, z in that order, against the
// squared near and mid radii; inactive slots are dropped (class 3).
__global__ void k_lod_classify(PBuf P, const uint8_t* __restrict__ active, uint32_t count, float px,
                               float py, float pz, float near2, float mid2,
                               uint8_t* __restrict__ cls, float* __restrict__ dist) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= count) return;
  if (!active[i]) {
    cls[i] = 3;
    return;
  }
  const float dx = P.x[0][i] - px;
  const float dy = P.x[1][i] - py;
  const float dz = P.x[2][i] - pz;
  const float d2 = dx * dx + dy * dy + dz * dz;
  dist[i] = d2;
  cls[i] = d2 < near2 ? 0 : (d2 < mid2 ? 1 : 2);
}

// Render filter (lib.rs:1900-1936): list entries that are still live
// (slot < count and active) keep class 0.
__global__ void k_live_filter(const uint32_t* __restrict__ list, uint32_t n,
                              const uint8_t* __restrict__ active, uint32_t count,
                              uint8_t* __restrict__ cls) {
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= n) return;
  const uint32_t s = list[k];
  cls[k] = (s < count && active[s]) ? 0 : 3;
}

__global__ __launch_bounds__(kSelThreads) void k_sel_count(const uint8_t* __restrict__ cls,
                                                           uint32_t n, uint32_t* __restrict__ bcnt) {
  __shared__ uint32_t c[kClasses];
  if (threadIdx.x < kClasses) c[threadIdx.x] = 0;
  __syncthreads();
  uint32_t loc[kClasses] = {0, 0, 0};
  const uint32_t base = blockIdx.x * kSelTile + threadIdx.x * kSelPer;
#pragma unroll
  for (int k = 0; k < kSelPer; k++) {
    const uint32_t i = base + k;
    if (i < n) {
      const uint8_t v = cls[i];
      loc[0] += v == 0;
      loc[1] += v == 1;
      loc[2] += v == 2;
    }
  }
#pragma unroll
  for (int q = 0; q < kClasses; q++)
    if (loc[q]) atomicAdd(&c[q], loc[q]);
  __syncthreads();
  if (threadIdx.x < kClasses) bcnt[blockIdx.x * kClasses + threadIdx.x] = c[threadIdx.x];
}

// Workgroup-wide inclusive scan of one value per thread (Hillis-Steele in
// LDS); `total` gets the workgroup sum.  Every thread of the workgroup calls it.
__device__ uint32_t wg_inclusive_scan(uint32_t v, uint32_t* s, uint32_t& total) {
  s[threadIdx.x] = v;
  __syncthreads();
  for (uint32_t d = 1; d < kSelThreads; d <<= 1) {
    const uint32_t add = threadIdx.x >= d ? s[threadIdx.x - d] : 0u;
    __syncthreads();
    s[threadIdx.x] += add;
    __syncthreads();
  }
  const uint32_t r = s[threadIdx.x];
  total = s[kSelThreads - 1];
  __syncthreads();
  return r;
}

// One workgroup: exclusive offset of every tile per class, and the totals.
__global__ __launch_bounds__(kSelThreads) void k_sel_scan(const uint32_t* __restrict__ bcnt,
                                                          uint32_t nblk, uint32_t* __restrict__ boff,
                                                          uint32_t* __restrict__ tot) {
  __shared__ uint32_t s[kSelThreads];
  for (int q = 0; q < kClasses; q++) {
    uint32_t run = 0;
    for (uint32_t b0 = 0; b0 < nblk; b0 += kSelThreads) {
      const uint32_t b = b0 + threadIdx.x;
      const uint32_t v = b < nblk ? bcnt[b * kClasses + q] : 0u;
      uint32_t chunk = 0;
      const uint32_t inc = wg_inclusive_scan(v, s, chunk);
      if (b < nblk) boff[b * kClasses + q] = run + inc - v;
      run += chunk;
    }
    if (threadIdx.x == 0) tot[q] = run;
  }
}

// Items of class q go to out_q in item order (src[i], or i without src);
// class-0 items also write their key's bits to key0 when given.
__global__ __launch_bounds__(kSelThreads) void k_sel_scatter(
    const uint8_t* __restrict__ cls, const uint32_t* __restrict__ src, const float* __restrict__ key,
    uint32_t n, const uint32_t* __restrict__ boff, uint32_t* __restrict__ out0,
    uint32_t* __restrict__ out1, uint32_t* __restrict__ out2, uint32_t* __restrict__ key0) {
  __shared__ uint32_t s[kSelThreads];
  const uint32_t base = blockIdx.x * kSelTile + threadIdx.x * kSelPer;
  uint8_t c[kSelPer];
  uint32_t loc[kClasses] = {0, 0, 0};
#pragma unroll
  for (int k = 0; k < kSelPer; k++) {
    const uint32_t i = base + k;
    c[k] = i < n ? cls[i] : (uint8_t)3;
    loc[0] += c[k] == 0;
    loc[1] += c[k] == 1;
    loc[2] += c[k] == 2;
  }
  uint32_t pos[kClasses];
#pragma unroll
  for (int q = 0; q < kClasses; q++) {
    uint32_t tot = 0;
    const uint32_t inc = wg_inclusive_scan(loc[q], s, tot);
    pos[q] = boff[blockIdx.x * kClasses + q] + inc - loc[q];
  }
#pragma unroll
  for (int k = 0; k < kSelPer; k++) {
    const uint32_t i = base + k;
    const uint8_t q = c[k];
    if (q >= kClasses) continue;
    const uint32_t item = src ? src[i] : i;
    if (q == 0) {
      if (out0) out0[pos[0]] = item;
      if (key0) key0[pos[0]] = __float_as_uint(key[i]);
      pos[0]++;
    } else if (q == 1) {
      if (out1) out1[pos[1]] = item;
      pos[1]++;
    } else {
      if (out2) out2[pos[2]] = item;
      pos[2]++;
    }
  }
}

// Far tier (simulation.rs:622-690, gravity and position part): f32
// v_y -= 9.81 far_dt; y += v_y far_dt, for every cached far slot.
__global__ void k_far_update(PBuf P, const uint32_t* __restrict__ far, uint32_t n, float far_dt) {
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= n) return;
  const uint32_t s = far[k];
  float v = P.v[1][s];
  v = v - 9.81f * far_dt;
  P.v[1][s] = v;
  P.x[1][s] = P.x[1][s] + v * far_dt;
}

__global__ void k_gather_pos(PBuf P, const uint32_t* __restrict__ list, uint32_t n,
                             float* __restrict__ out3) {
  const uint32_t k = blockIdx.x * 256u + threadIdx.x;
  if (k >= n) return;
  const uint32_t s = list[k];
#pragma unroll
  for (int a = 0; a < 3; a++) out3[(size_t)k * 3 + a] = P.x[a][s];
}

}  // namespace

uint32_t select_tiles(uint32_t n) { return (n + kSelTile - 1) / kSelTile; }

void launch_lod_classify(hipStream_t st, PBuf P, const uint8_t* active, uint32_t count,
                         const float player[3], float near2, float mid2, uint8_t* cls,
                         float* dist) {
  if (count)
    k_lod_classify<<<grid_of(count), 256, 0, st>>>(P, active, count, player[0], player[1],
                                                   player[2], near2, mid2, cls, dist);
}

void launch_live_filter(hipStream_t st, const uint32_t* list, uint32_t n, const uint8_t* active,
                        uint32_t count, uint8_t* cls) {
  if (n) k_live_filter<<<grid_of(n), 256, 0, st>>>(list, n, active, count, cls);
}

void launch_select(hipStream_t st, const uint8_t* cls, const uint32_t* src, const float* key,
                   uint32_t n, uint32_t* bcnt, uint32_t* boff, uint32_t* tot, uint32_t* out0,
                   uint32_t* out1, uint32_t* out2, uint32_t* key0) {
  const uint32_t nblk = select_tiles(n);
  if (nblk) k_sel_count<<<nblk, kSelThreads, 0, st>>>(cls, n, bcnt);
  k_sel_scan<<<1, kSelThreads, 0, st>>>(bcnt, nblk, boff, tot);
  if (nblk)
    k_sel_scatter<<<nblk, kSelThreads, 0, st>>>(cls, src, key, n, boff, out0, out1, out2, key0);
}

size_t sort_pairs_temp_bytes(uint32_t n) {
  size_t bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr,
                                     (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)n, 0, 32);
  return bytes;
}

hipError_t launch_sort_pairs(hipStream_t st, void* tmp, size_t tmp_bytes, const uint32_t* kin,
                             uint32_t* kout, const uint32_t* vin, uint32_t* vout, uint32_t n) {
  // LSD radix sort: stable, so equal keys keep their input order.
  return hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kin, kout, vin, vout, (int)n, 0, 32,
                                            st);
}

size_t sort_pairs_u64_temp_bytes(uint32_t n, int end_bit) {
  size_t bytes = 0;
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const unsigned long long*)nullptr,
                                           (unsigned long long*)nullptr, (const uint32_t*)nullptr,
                                           (uint32_t*)nullptr, (int)n, 0, end_bit);
  return bytes;
}

void launch_sort_pairs_u64(hipStream_t st, void* tmp, size_t tmp_bytes,
                           const unsigned long long* kin, unsigned long long* kout,
                           const uint32_t* vin, uint32_t* vout, uint32_t n, int end_bit) {
  if (n)
    (void)hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kin, kout, vin, vout, (int)n, 0,
                                             end_bit, st);
}

void launch_far_update(hipStream_t st, PBuf P, const uint32_t* far, uint32_t n, float far_dt) {
  if (n) k_far_update<<<grid_of(n), 256, 0, st>>>(P, far, n, far_dt);
}

void launch_gather_pos(hipStream_t st, PBuf P, const uint32_t* list, uint32_t n, float* out3) {
  if (n) k_gather_pos<<<grid_of(n), 256, 0, st>>>(P, list, n, out3);
}

}  // namespace aemh
