// aemh_launch.h -- host-callable launchers for the kernels of each
// translation unit.  All launches go to the caller's stream and are
// asynchronous; grid sizes depend only on sizes the host knows, so a whole
// substep can be captured in a hipGraph.
#pragma once
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "aemh_common.h"

namespace aemh {

// Per-kernel timing of profiled steps.  While a stage is timed, its launches
// hand the stage's events to hipExtLaunchKernel: the start comes from the
// first kernel's dispatch and the stop from the last one's, so no marker
// packets sit between the kernels and the timed sequence is the one the
// graph replays.
struct StageEvents {
  hipEvent_t start = nullptr, stop = nullptr;
  bool launched = false;
};
extern thread_local StageEvents g_stage;

template <typename... KArgs, typename... Args>
inline void launch_k(void (*k)(KArgs...), dim3 grid, dim3 block, size_t shmem, hipStream_t st,
                     Args... args) {
  static_assert(sizeof...(KArgs) == sizeof...(Args), "kernel argument count");
  if (g_stage.stop) {
    hipExtLaunchKernelGGL(k, grid, block, (uint32_t)shmem, st,
                          g_stage.launched ? nullptr : g_stage.start, g_stage.stop, 0u,
                          static_cast<KArgs>(args)...);
    g_stage.launched = true;
  } else {
    k<<<grid, block, shmem, st>>>(static_cast<KArgs>(args)...);
  }
}

// Launch parameters the autotuner chooses (defaults: the 1M-particle snow
// workload's best).  P2G always runs 256 particles per workgroup.
struct LaunchShape {
  int p2g_tiles = 3;     // LDS tiles per P2G workgroup: 2 to 5
  int p2g_partials = 0;  // P2G tile flush: 0 one HBM atomic per node value,
                         // 1 partial stores gathered by the grid update
  int p2g_pairs = 0;     // FP32 mode: 1 pairs lanes of one base cell and adds
                         // their node values across the pair before the LDS atomics
  int p2g_layers = 0;    // 1 re-assigns lanes by rank within their base cell
                         // so one LDS atomic's lanes hit distinct nodes
  int g2p_wg = 256;      // G2P workgroup size: 64, 128, 256 or 512
  int upd_wgs = 2048;    // persistent grid-update / clear workgroups
};

// ---- k_common.hip --------------------------------------------------------
void launch_count(hipStream_t st, PBuf cur, const SceneDev* sc, int ns, uint32_t n, SortDev s,
                  uint32_t nb_total);
// Scan of the block counts (with the halo activation flags).
void launch_scan(hipStream_t st, const SceneDev* sc, int ns, SortDev s, uint32_t nb);
// Block-order permutation, active-block list and the zeroing of its blocks'
// accumulators (acc: the plane layout, kAcc*).
void launch_scatter(hipStream_t st, uint32_t n, uint32_t nb, SortDev s, GridDev g, int acc);
void launch_clear(hipStream_t st, SortDev s, GridDev g, int acc, int upd_wgs);
// canonical (capacity) <-> working set
void launch_gather_work(hipStream_t st, PBuf canon, PBuf work, const uint32_t* idx, uint32_t n,
                        int storage);
void launch_scatter_work(hipStream_t st, PBuf work, PBuf canon, const uint32_t* idx, uint32_t n,
                         int storage);
void launch_fill(hipStream_t st, float* p, float v, uint32_t n);
// Dense x-fastest readback of one scene's grid (active blocks only).
void launch_grid_readback(hipStream_t st, const SceneDev* sc_d, int scene, SceneDev sc,
                          SortDev s, GridDev g, int acc, float* m, float* p3, float* f3,
                          float* v3, long long* q7, int upd_wgs);
// FP64 sums over the working set: mass, momentum xyz, kinetic energy.
// out5 needs room for kConsScratch doubles: the 5 sums, then the 5 partial
// sums of each of the kConsWgs reduction workgroups (every one is written).
constexpr int kConsWgs = 512;
constexpr int kConsScratch = 5 + 5 * kConsWgs;
void launch_conservation(hipStream_t st, PBuf cur, uint32_t n, double* out5);

// ---- k_mpss.hip: resident MpssBuffer lifecycle ------------------------------
void launch_mpss_reset(hipStream_t st, PBuf P, MpssDev L, const uint32_t* slots, uint32_t n);
void launch_mpss_set_active(hipStream_t st, uint8_t* act, const uint32_t* slots, uint32_t n,
                            uint8_t v);
void launch_mpss_gather(hipStream_t st, PBuf S, MpssDev LS, PBuf D, MpssDev LD,
                        const uint32_t* perm, uint32_t cap, uint32_t count);
void launch_mpss_age(hipStream_t st, MpssDev L, uint32_t count, float dt, uint32_t* out,
                     uint32_t* nout);
void launch_rows_to_slots(hipStream_t st, const float* rows, Ptr9 dst, int comps,
                          const uint32_t* slots, uint32_t n);
void launch_slots_to_rows(hipStream_t st, Ptr9 src, float* rows, int comps, const uint32_t* slots,
                          uint32_t n);
void launch_u16_to_slots(hipStream_t st, const uint16_t* v, uint16_t* dst, const uint32_t* slots,
                         uint32_t n);
void launch_slots_to_u16(hipStream_t st, const uint16_t* src, uint16_t* v, const uint32_t* slots,
                         uint32_t n);

// ---- k_tier.hip: LOD tier orchestrator ------------------------------------
// Tiles of the ordered compaction (bcnt / boff hold 3 words per tile).
uint32_t select_tiles(uint32_t n);
void launch_lod_classify(hipStream_t st, PBuf P, const uint8_t* active, uint32_t count,
                         const float player[3], float near2, float mid2, uint8_t* cls,
                         float* dist);
void launch_live_filter(hipStream_t st, const uint32_t* list, uint32_t n, const uint8_t* active,
                        uint32_t count, uint8_t* cls);
// Items of class 0/1/2 (cls, one byte per item) to out0/out1/out2 in item
// order; tot[3] gets the class sizes.  Null outputs are skipped.
void launch_select(hipStream_t st, const uint8_t* cls, const uint32_t* src, const float* key,
                   uint32_t n, uint32_t* bcnt, uint32_t* boff, uint32_t* tot, uint32_t* out0,
                   uint32_t* out1, uint32_t* out2, uint32_t* key0);
size_t sort_pairs_temp_bytes(uint32_t n);
// Radix sort of 64-bit keys (bits [0, end_bit)) with 32-bit values.
size_t sort_pairs_u64_temp_bytes(uint32_t n, int end_bit);
void launch_sort_pairs_u64(hipStream_t st, void* tmp, size_t tmp_bytes,
                           const unsigned long long* kin, unsigned long long* kout,
                           const uint32_t* vin, uint32_t* vout, uint32_t n, int end_bit);
hipError_t launch_sort_pairs(hipStream_t st, void* tmp, size_t tmp_bytes, const uint32_t* kin,
                             uint32_t* kout, const uint32_t* vin, uint32_t* vout, uint32_t n);
void launch_far_update(hipStream_t st, PBuf P, const uint32_t* far, uint32_t n, float far_dt);
void launch_gather_pos(hipStream_t st, PBuf P, const uint32_t* list, uint32_t n, float* out3);

// ---- k_fast.hip (FP32 mode) and k_det.hip (deterministic mode) -----------
#define AEMH_DECLARE_MODE(NS)                                                               \
  namespace NS {                                                                            \
  void launch_p2g(hipStream_t st, const LaunchShape& ls, PBuf cur, const SceneDev* sc,      \
                  int ns, const MatDev* mats, int nmat, uint32_t n, SortDev s, GridDev g,   \
                  int storage);                                                             \
  void launch_grid_update(hipStream_t st, const LaunchShape& ls, const SceneDev* sc,        \
                          int ns, SortDev s, GridDev g, uint32_t n, uint32_t nb_total);     \
  void launch_g2p(hipStream_t st, const LaunchShape& ls, PBuf cur, PBuf nxt,                \
                  const SceneDev* sc, int ns, const MatDev* mats, int nmat, uint32_t n,     \
                  SortDev s, GridDev g, int storage, int inplace, int count);               \
  }
AEMH_DECLARE_MODE(fast)
AEMH_DECLARE_MODE(det)
#undef AEMH_DECLARE_MODE
namespace fast {
// The reference-order mode's grid update on 7 FP32 planes (m, p, f).
void launch_grid_update_mpf(hipStream_t st, const LaunchShape& ls, const SceneDev* sc, int ns,
                            SortDev s, GridDev g, uint32_t n, uint32_t nb_total);
}
namespace det {
// Reference-order FP32 P2G: the grid's FP32 sums (accf) of the active
// blocks, each node's contributions added in list order.
void launch_p2g_serial(hipStream_t st, PBuf cur, const SceneDev* sc, int ns, const MatDev* mats,
                       int nmat, uint32_t n, SerialDev sd, GridDev g, int storage);
}

}  // namespace aemh
