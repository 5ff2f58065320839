// k_mpss.hip -- device side of the resident MpssBuffer lifecycle
// (ae_particle/src/mpss.rs:189-297, 431-445): slot resets on spawn, active
// flags, the compaction gather and lifetime ageing.  The slot bookkeeping
// (free list, count) is host state of the context; these kernels apply it to
// the particle arrays that live in HBM.
#include <float.h>

#include "aemh_dev.h"
#include "aemh_launch.h"

namespace aemh {
namespace {

uint32_t grid_of(uint32_t n) { return (n + 255u) / 256u; }

// MpssBuffer::reset_slot (mpss.rs:215-232) + active: pos, vel 0; strain I;
// jacobian 1; mass 1; material 0; lifetime f32::MAX; age 0.  C is not reset
// (reset_slot leaves it); J_p, the snow model's plastic ratio, restarts at 1.
__global__ void k_mpss_reset(PBuf P, MpssDev L, const uint32_t* __restrict__ slots, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const uint32_t s = slots[i];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    P.x[a][s] = 0.0f;
    P.v[a][s] = 0.0f;
  }
#pragma unroll
  for (int k = 0; k < 9; k++) P.F[k][s] = (k % 4 == 0) ? 1.0f : 0.0f;
  P.J[s] = 1.0f;
  P.Jp[s] = 1.0f;
  P.mass[s] = 1.0f;
  P.mat[s] = 0;
  L.lifetime[s] = FLT_MAX;
  L.age[s] = 0.0f;
  L.active[s] = 1;
}

__global__ void k_mpss_set_active(uint8_t* act, const uint32_t* __restrict__ slots, uint32_t n,
                                  uint8_t v) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) act[slots[i]] = v;
}

// Compaction: slot i takes the contents of slot perm[i] (the permutation the
// reference's swap sequence produces); the first `count` slots are active.
__global__ void k_mpss_gather(PBuf S, MpssDev LS, PBuf D, MpssDev LD,
                              const uint32_t* __restrict__ perm, uint32_t cap, uint32_t count) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= cap) return;
  const uint32_t p = perm[i];
#pragma unroll
  for (int a = 0; a < 3; a++) {
    D.x[a][i] = S.x[a][p];
    D.v[a][i] = S.v[a][p];
  }
#pragma unroll
  for (int k = 0; k < 9; k++) {
    D.C[k][i] = S.C[k][p];
    D.F[k][i] = S.F[k][p];
  }
  D.J[i] = S.J[p];
  D.Jp[i] = S.Jp[p];
  D.mass[i] = S.mass[p];
  D.mat[i] = S.mat[p];
  LD.lifetime[i] = LS.lifetime[p];
  LD.age[i] = LS.age[p];
  LD.active[i] = i < count ? 1 : 0;
}

// update_lifetimes (mpss.rs:431-445): slots below `count` that are active age
// by dt; those reaching their lifetime are listed (any order; the host sorts).
__global__ void k_mpss_age(MpssDev L, uint32_t count, float dt, uint32_t* out, uint32_t* nout) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= count || !L.active[i]) return;
  const float a = L.age[i] + dt;
  L.age[i] = a;
  if (a >= L.lifetime[i]) out[atomicAdd(nout, 1u)] = i;
}

// Host [n][comps] rows (staged in HBM) into / out of slots of SoA arrays.
__global__ void k_rows_to_slots(const float* __restrict__ rows, Ptr9 dst, int comps,
                                const uint32_t* __restrict__ slots, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const uint32_t s = slots[i];
  for (int c = 0; c < comps; c++) dst.p[c][s] = rows[(size_t)i * comps + c];
}
__global__ void k_slots_to_rows(Ptr9 src, float* __restrict__ rows, int comps,
                                const uint32_t* __restrict__ slots, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  const uint32_t s = slots[i];
  for (int c = 0; c < comps; c++) rows[(size_t)i * comps + c] = src.p[c][s];
}
__global__ void k_u16_to_slots(const uint16_t* __restrict__ v, uint16_t* dst,
                               const uint32_t* __restrict__ slots, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) dst[slots[i]] = v[i];
}
__global__ void k_slots_to_u16(const uint16_t* src, uint16_t* __restrict__ v,
                               const uint32_t* __restrict__ slots, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i < n) v[i] = src[slots[i]];
}

}  // namespace

void launch_mpss_reset(hipStream_t st, PBuf P, MpssDev L, const uint32_t* slots, uint32_t n) {
  if (n) k_mpss_reset<<<grid_of(n), 256, 0, st>>>(P, L, slots, n);
}
void launch_mpss_set_active(hipStream_t st, uint8_t* act, const uint32_t* slots, uint32_t n,
                            uint8_t v) {
  if (n) k_mpss_set_active<<<grid_of(n), 256, 0, st>>>(act, slots, n, v);
}
void launch_mpss_gather(hipStream_t st, PBuf S, MpssDev LS, PBuf D, MpssDev LD,
                        const uint32_t* perm, uint32_t cap, uint32_t count) {
  if (cap) k_mpss_gather<<<grid_of(cap), 256, 0, st>>>(S, LS, D, LD, perm, cap, count);
}
void launch_mpss_age(hipStream_t st, MpssDev L, uint32_t count, float dt, uint32_t* out,
                     uint32_t* nout) {
  if (count) k_mpss_age<<<grid_of(count), 256, 0, st>>>(L, count, dt, out, nout);
}
void launch_rows_to_slots(hipStream_t st, const float* rows, Ptr9 dst, int comps,
                          const uint32_t* slots, uint32_t n) {
  if (n) k_rows_to_slots<<<grid_of(n), 256, 0, st>>>(rows, dst, comps, slots, n);
}
void launch_slots_to_rows(hipStream_t st, Ptr9 src, float* rows, int comps, const uint32_t* slots,
                          uint32_t n) {
  if (n) k_slots_to_rows<<<grid_of(n), 256, 0, st>>>(src, rows, comps, slots, n);
}
void launch_u16_to_slots(hipStream_t st, const uint16_t* v, uint16_t* dst, const uint32_t* slots,
                         uint32_t n) {
  if (n) k_u16_to_slots<<<grid_of(n), 256, 0, st>>>(v, dst, slots, n);
}
void launch_slots_to_u16(hipStream_t st, const uint16_t* src, uint16_t* v, const uint32_t* slots,
                         uint32_t n) {
  if (n) k_slots_to_u16<<<grid_of(n), 256, 0, st>>>(src, v, slots, n);
}

}  // namespace aemh
