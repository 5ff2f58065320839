// aemh_api.hip -- host side of the C ABI declared in include/aemh.h.
//
// Owns every device buffer of a context: the canonical MpssBuffer-shaped
// particle arrays (capacity n, original slot order), two working buffers of
// the active set that G2P ping-pongs between (leaving them sorted by base
// block), the block-major grid, the sort bookkeeping, and the scene table.
// A substep is a fixed sequence of kernels on one stream; repeated substeps
// are captured once into a hipGraph per ping-pong parity and replayed.
#pragma clang fp contract(off)

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <float.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <thread>
#include <string>
#include <vector>

#include "../../include/aemh.h"
#include "aemh_common.h"
#include "aemh_launch.h"

using namespace aemh;

namespace {

const char* kVersion = "aemh 0.1 (gfx950)";

// G2P re-sorts the working set by block when the count kernel's wave runs
// (one per change of block along the slots) have grown by a quarter over
// the count right after the last re-sort, and at least every kResortMax
// substeps; the other substeps update the particles in place.
constexpr uint32_t kResortMax = 64;
// When both copies of the working set fit in MI355X's 256 MB Infinity Cache
// the re-sorting G2P's extra traffic (mass, material, id, perm) stays in
// that cache, and every kResortResident-th substep re-sorts (on the 1M-particle
// snow workload, every 2nd: step kernels 0.207 against 0.209-0.217 ms re-sorting
// every substep, three runs each); above it, in-place steps save HBM traffic.
constexpr size_t kInfinityCacheBytes = (size_t)256 << 20;
constexpr uint32_t kResortResident = 2;
// Working-set slots are addressed with 32-bit byte offsets in the kernels.
constexpr uint32_t kMaxWork = 1u << 30;

// Device memory of aemh_conservation_f64_noalloc: the 5 sums followed by the
// 5 partial sums of each reduction workgroup (kConsScratch doubles, 20.5 KB),
// one copy per device in the loaded code object, held until the process
// ends.  In host code the name is not a device address: the kernels get the
// address hipGetSymbolAddress returns for the current device (kept per
// device in g_scratch by scratch_for_device, called by aemh_create).
std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};

struct ParticleStore {
  uint32_t cap = 0;
  int storage = 0;
  void* slab = nullptr;
  PBuf pb{};
};

struct ProfRec {
  double total_ms = 0;
  uint64_t count = 0;
};

}  // namespace

static __device__ __align__(256) double scratch[aemh::kConsScratch];

struct aemh_ctx {
  int device = 0;
  // this device's address of `scratch` (nullptr if it could not be had)
  double* cons_scratch = nullptr;
  std::string scratch_err;
  hipStream_t stream = nullptr;
  std::string err;
  aemh_config cfg{};
  // context grid (single-scene API)
  uint32_t nx = 0, ny = 0, nz = 0;
  float dx = 0, origin[3] = {0, 0, 0};
  bool grid_set = false;
  // materials (n + 1 entries: the last is the unwrap_or fallback)
  std::vector<MatDev> mats_h;
  MatDev* mats_d = nullptr;
  uint32_t nmat = 0;
  // particles
  ParticleStore canon;
  ParticleStore work[2];
  int cur = 0;
  uint32_t n_work = 0;
  bool work_valid = false, work_dirty = false;
  std::vector<uint32_t> idx_h;
  uint32_t* idx_d = nullptr;
  float* staging = nullptr;
  size_t staging_bytes = 0;
  // resident MpssBuffer lifecycle (aemh_mpss_reserve): device lifecycle
  // arrays of the canonical store, host bookkeeping mirroring mpss.rs
  bool mpss = false;
  MpssDev life{};
  MpssDev life_tmp{};
  ParticleStore canon_tmp;
  std::vector<uint8_t> act_h;
  std::vector<uint32_t> free_list;
  uint32_t mpss_count = 0;
  uint32_t* slots_d = nullptr;  // slot lists / permutation for lifecycle kernels
  size_t slots_cap = 0;
  uint32_t* expired_d = nullptr;  // [1 + cap]: count, then expired slots
  bool idx_explicit = false;      // the caller fixed the list (aemh_set_indices)
  bool work_default = false;      // the work set is the default set below
  size_t idx_cap = 0;
  // LOD tier orchestrator (aemh_tier_tick): counters and cached lists of
  // SimulationManager (simulation.rs:97-103, 480-560), lists on device
  aemh_lod_config lod{};
  bool lod_set = false;
  uint32_t field_counter = 0, lod_counter = 0;
  float cached_player[3] = {0.0f, 0.0f, 0.0f};
  uint32_t tier_n[3] = {0, 0, 0};  // cached near / mid / far list sizes
  uint32_t tier_version = 0;       // bumped by each reclassification
  uint64_t work_tier_key = 0;      // tier lists the work set holds (0: none)
  uint32_t tier_cap = 0;
  uint32_t* tier_list[3] = {nullptr, nullptr, nullptr};
  uint8_t* tier_cls = nullptr;
  float* tier_dist = nullptr;
  uint32_t* tier_keys[2] = {nullptr, nullptr};
  uint32_t* tier_vals = nullptr;
  uint32_t* tier_bcnt = nullptr;  // [tiles][3] counts, then [tiles][3] offsets, then totals
  void* tier_sort_tmp = nullptr;
  size_t tier_sort_bytes = 0;
  // scenes
  std::vector<aemh_scene> scenes_h;  // batch table as given
  bool batch_mode = false;
  std::vector<SceneDev> scd_h;
  SceneDev* scd_d = nullptr;
  size_t scd_cap = 0;
  bool scd_valid = false;  // scd_d holds scd_h
  uint32_t nb_total = 0;
  // grid + sort
  GridDev grid{};
  int grid_q32 = 0;
  uint32_t grid_nb_cap = 0;
  SortDev sort{};
  uint32_t sort_n_cap = 0, sort_nb_cap = 0;
  size_t part_wg_cap = 0;
  int part_q32 = 0;
  LaunchShape shape;
  bool tuned = false;
  float tune_ms[3] = {0.0f, 0.0f, 0.0f};  // best P2G, grid update, G2P time
  // graphs: one per (ping-pong parity, re-sort step, count kernel or not),
  // built on first use
  hipGraphExec_t gexec[8] = {};
  bool graph_valid = false;
  // The last substep's G2P counted the sort of work[cur] as it stands (keys,
  // ranks, block counts and wave runs), so the next substep skips the count
  // kernel.  Cleared, with the counts, by any change made between substeps.
  bool counted = false;
  // in-place substeps between re-sorting ones
  uint32_t since_sort = 0;
  bool need_sort = true;
  bool fresh_sort = false;     // the last substep re-sorted
  bool unsynced = false;       // substeps enqueued by aemh_step_async, not yet synchronized
  uint32_t runs_base = 0;      // wave runs counted right after a re-sort
  uint32_t* runs_h = nullptr;  // host-mapped: wave runs of the last count
  bool use_graph = true;
  // profiling
  bool prof = false;
  std::map<std::string, ProfRec> prof_rec;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
  std::vector<hipEvent_t> ev_pool;
  // pinned readback word for overflow checks
  uint32_t* flag_h = nullptr;
  // reference-order FP32 mode (AEMH_ACCUM_F32_SERIAL): entry buffers for
  // ser_cap entries and the device copy of the host logf table
  SerialDev ser{};
  size_t ser_cap = 0;
  float* lntab_d = nullptr;
};

namespace {

// The grid's accumulator plane layout (kAcc*) of the configured mode.
static_assert(kAccF32Fold == AEMH_ACCUM_F32 && kAccQ32 == AEMH_ACCUM_Q32_32 &&
                  kAccF32Mpf == AEMH_ACCUM_F32_SERIAL,
              "plane layouts are numbered as the accumulation modes");
int acc_layout(const aemh_ctx* c) { return (int)c->cfg.accum; }

// This device's address of `scratch`, looked up once per device with the
// device current (each device holds its own copy at its own address).
hipError_t scratch_for_device(int device, double** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = static_cast<double*>(p);
  }
  *out = g_scratch[device];
  return hipSuccess;
}

aemh_status fail(aemh_ctx* c, aemh_status s, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (c) c->err = buf;
  return s;
}

#define HIPTRY(ctx, expr)                                                                    \
  do {                                                                                       \
    hipError_t e_ = (expr);                                                                  \
    if (e_ != hipSuccess) {                                                                  \
      return fail(ctx, e_ == hipErrorOutOfMemory ? AEMH_ERR_OOM : AEMH_ERR_HIP, "%s: %s (%s)", \
                  #expr, hipGetErrorName(e_), hipGetErrorString(e_));                        \
    }                                                                                        \
  } while (0)

#define CHECK_LAUNCH(ctx)                                                                    \
  do {                                                                                       \
    hipError_t e_ = hipGetLastError();                                                       \
    if (e_ != hipSuccess)                                                                    \
      return fail(ctx, AEMH_ERR_HIP, "kernel launch: %s (%s)", hipGetErrorName(e_),          \
                  hipGetErrorString(e_));                                                    \
  } while (0)

// Per-slot arrays are carved from one slab, 256 B aligned, each followed by
// an odd multiple of 256 B: with power-of-two particle counts the arrays
// would otherwise start power-of-two bytes apart, and the ~26 streams a
// kernel reads or writes at the same slot would meet in the same HBM
// channels.
constexpr size_t kPlanePad = 9 * 256;

template <typename T>
T* carve(char*& p, size_t count) {
  T* r = reinterpret_cast<T*>(p);
  size_t bytes = (count * sizeof(T) + 255) & ~(size_t)255;
  p += bytes + kPlanePad;
  return r;
}

void carve_store(char* p, PBuf& pb, uint32_t n, int storage) {
  for (int a = 0; a < 3; a++) pb.x[a] = carve<float>(p, n);
  for (int a = 0; a < 3; a++) pb.v[a] = carve<float>(p, n);
  for (int k = 0; k < 9; k++) {
    if (storage == 0) {
      pb.C[k] = carve<float>(p, n);
      pb.F[k] = carve<float>(p, n);
    } else {
      pb.CF[k] = carve<uint32_t>(p, n);
    }
  }
  pb.J = carve<float>(p, n);
  pb.Jp = carve<float>(p, n);
  pb.mass = carve<float>(p, n);
  pb.mat = carve<uint16_t>(p, n);
  pb.id = carve<uint32_t>(p, n);
}

size_t store_bytes(uint32_t n, int storage) {
  PBuf pb{};
  char* base = reinterpret_cast<char*>((uintptr_t)4096);  // layout only, never dereferenced
  carve_store(base, pb, n, storage);
  return (size_t)((char*)pb.id - base) + (((size_t)n * 4 + 255) & ~(size_t)255) + kPlanePad;
}

// Both copies of the working set fit in the Infinity Cache.  Such sets
// re-sort every kResortResident-th substep, and their G2P counts the next
// substep's sort: on the 1M-particle snow workload the count kernel's 9.8 us
// go for 5.4 us more G2P.  On larger sets G2P grows by as much as the count
// kernel takes (16M particles: 93 us against 97 us), and they keep it.
bool cache_resident(const aemh_ctx* c) {
  return 2 * store_bytes(c->n_work, c->work[0].storage) <= kInfinityCacheBytes;
}

aemh_status alloc_store(aemh_ctx* c, ParticleStore& s, uint32_t n, int storage) {
  if (s.slab && s.cap >= n && s.storage == storage) return AEMH_OK;
  if (s.slab) {
    hipFree(s.slab);
    s.slab = nullptr;
  }
  s.cap = n;
  s.storage = storage;
  s.pb = PBuf{};
  if (n == 0) return AEMH_OK;
  HIPTRY(c, hipMalloc(&s.slab, store_bytes(n, storage)));
  carve_store((char*)s.slab, s.pb, n, storage);
  return AEMH_OK;
}

aemh_status ensure_staging(aemh_ctx* c, size_t bytes) {
  if (c->staging_bytes >= bytes) return AEMH_OK;
  if (c->staging) hipFree(c->staging);
  c->staging = nullptr;
  c->staging_bytes = 0;
  HIPTRY(c, hipMalloc(&c->staging, bytes));
  c->staging_bytes = bytes;
  return AEMH_OK;
}

// The block counts and wave runs G2P left for the next substep are zeroed
// (stream-ordered after it) and that substep runs the count kernel.
void drop_counts(aemh_ctx* c) {
  if (!c->counted) return;
  c->counted = false;
  hipError_t e = hipMemsetAsync(c->sort.cnt, 0, (size_t)c->sort_nb_cap * 4, c->stream);
  if (e == hipSuccess)
    e = hipMemsetAsync(c->sort.runs, 0, (size_t)kRunLanes * kRunStride * 4, c->stream);
  if (e != hipSuccess) c->err = std::string("clearing the sort counts: ") + hipGetErrorName(e);
}

// Every change that rebuilds the graphs (buffers, launch shape, working set,
// configuration, materials) also drops the counts G2P made.
void invalidate_graphs(aemh_ctx* c) {
  for (int i = 0; i < 8; i++)
    if (c->gexec[i]) {
      hipGraphExecDestroy(c->gexec[i]);
      c->gexec[i] = nullptr;
    }
  c->graph_valid = false;
  drop_counts(c);
}

// ---------------------------------------------------------------------------
// Material constants with the reference's f32 evaluation order
// (constitutive.rs:55-82, :485-490); contraction is off in this file.
// ---------------------------------------------------------------------------
MatDev make_mat(const float* p, int model, const float* ext) {
  MatDev m{};
  const float E = p[0];
  float nu = p[1];
  if (nu < -0.999f) nu = -0.999f;
  if (nu > 0.4999f) nu = 0.4999f;
  m.lambda = E * nu / ((1.0f + nu) * (1.0f - 2.0f * nu));
  m.mu = E / (2.0f * (1.0f + nu));
  m.bulk = E / (3.0f * (1.0f - 2.0f * nu));
  m.yield_stress = p[2];
  m.density = p[4];
  const float rads_per_deg = 3.14159265358979323846f / 180.0f;
  const float phi = p[5] * rads_per_deg;
  const float sphi = sinf(phi), cphi = cosf(phi);
  const float sqrt3 = sqrtf(3.0f);
  m.dp_alpha = (2.0f * sphi) / (sqrt3 * (3.0f - sphi));
  m.dp_k = (6.0f * p[6] * cphi) / (sqrt3 * (3.0f - sphi));
  m.theta_c = ext ? ext[0] : 2.5e-2f;
  m.theta_s = ext ? ext[1] : 7.5e-3f;
  m.xi = ext ? ext[2] : 10.0f;
  m.model = model;
  return m;
}

SceneDev make_scene(const aemh_scene& s, const aemh_config& cfg, uint32_t block_off) {
  SceneDev d{};
  d.first = s.first;
  d.count = s.count;
  d.n[0] = (int)s.nx;
  d.n[1] = (int)s.ny;
  d.n[2] = (int)s.nz;
  for (int a = 0; a < 3; a++) d.nb[a] = (d.n[a] + 3) / 4;
  d.block_off = block_off;
  d.dx = s.dx;
  d.inv_dx = 1.0f / s.dx;  // Grid::new (mpm_solver.rs:100)
  for (int a = 0; a < 3; a++) d.o[a] = s.origin[a];
  d.dt = s.dt;
  for (int a = 0; a < 3; a++) d.g[a] = cfg.gravity[a];
  d.damping = cfg.damping;
  d.bc_margin = (int)cfg.bc_margin;
  d.bc_rest = cfg.bc_restitution;
  d.vel_only = s.velocity_only ? 1u : 0u;
  uint32_t cm = s.clamp_mode;
  d.window_clamp = (cm == AEMH_CLAMP_WINDOW || (cm == AEMH_CLAMP_AUTO && s.count >= 50000)) ? 1u : 0u;
  d.gmax_x = (float)(s.nx - 1) * s.dx;
  d.gmax_z = (float)(s.nz - 1) * s.dx;
  d.four_inv_dx2 = 4.0f * d.inv_dx * d.inv_dx;
  return d;
}

// P2G tile partials (partial-flush mode only, allocated on first use): room
// for the largest tile budget, 216 node records of 8 values per tile, per
// 256-particle workgroup.
aemh_status ensure_partials(aemh_ctx* c) {
  const int q32 = c->cfg.accum == AEMH_ACCUM_Q32_32;
  const size_t wgs = ((size_t)(c->n_work ? c->n_work : 1) + 255) / 256;
  if (wgs <= c->part_wg_cap && q32 == c->part_q32 && (c->grid.partf || c->grid.partq))
    return AEMH_OK;
  if (c->grid.partf) hipFree(c->grid.partf);
  if (c->grid.partq) hipFree(c->grid.partq);
  c->grid.partf = nullptr;
  c->grid.partq = nullptr;
  c->part_wg_cap = 0;
  const size_t vals = wgs * kMaxTileBudget * (size_t)kTileNodes * kDetFields;
  if (q32)
    HIPTRY(c, hipMalloc(&c->grid.partq, vals * sizeof(long long)));
  else
    HIPTRY(c, hipMalloc(&c->grid.partf, vals * sizeof(float)));
  c->part_wg_cap = wgs;
  c->part_q32 = q32;
  invalidate_graphs(c);
  return AEMH_OK;
}

aemh_status ensure_grid_sort(aemh_ctx* c) {
  const uint32_t nb = c->nb_total;
  const uint32_t n = c->n_work;
  const int q32 = c->cfg.accum == AEMH_ACCUM_Q32_32;
  if (nb > c->grid_nb_cap || q32 != c->grid_q32 || !c->grid.gv) {
    if (c->grid.accf) hipFree(c->grid.accf);
    if (c->grid.accq) hipFree(c->grid.accq);
    if (c->grid.gv) hipFree(c->grid.gv);
    if (c->grid.partf) hipFree(c->grid.partf);
    if (c->grid.partq) hipFree(c->grid.partq);
    c->grid = GridDev{};
    c->part_wg_cap = 0;
    const size_t nodes = (size_t)nb * kBlockNodes;
    if (nodes >= (1ull << 32)) return fail(c, AEMH_ERR_ARG, "grid too large: %zu nodes", nodes);
    if (q32) {
      HIPTRY(c, hipMalloc(&c->grid.accq, kDetFields * nodes * sizeof(long long)));
      HIPTRY(c, hipMemset(c->grid.accq, 0, kDetFields * nodes * sizeof(long long)));
    } else {
      HIPTRY(c, hipMalloc(&c->grid.accf, 7 * nodes * sizeof(float)));
      HIPTRY(c, hipMemset(c->grid.accf, 0, 7 * nodes * sizeof(float)));
    }
    HIPTRY(c, hipMalloc(&c->grid.gv, nodes * sizeof(float4)));
    HIPTRY(c, hipMemset(c->grid.gv, 0, nodes * sizeof(float4)));
    c->grid.nodes = (uint32_t)nodes;
    c->grid_nb_cap = nb;
    c->grid_q32 = q32;
    invalidate_graphs(c);
  }
  if (nb > c->sort_nb_cap || !c->sort.cnt) {
    for (void* p : {(void*)c->sort.cnt, (void*)c->sort.off, (void*)c->sort.tsum, (void*)c->sort.traw,
                    (void*)c->sort.flag, (void*)c->sort.active, (void*)c->sort.bslot})
      if (p) hipFree(p);
    const uint32_t tiles = (nb + kScanTile - 1) / kScanTile;
    HIPTRY(c, hipMalloc(&c->sort.cnt, (size_t)nb * 4));
    HIPTRY(c, hipMemset(c->sort.cnt, 0, (size_t)nb * 4));
    HIPTRY(c, hipMalloc(&c->sort.off, (size_t)nb * 4));
    HIPTRY(c, hipMalloc(&c->sort.tsum, (size_t)(tiles + 1) * 4));
    HIPTRY(c, hipMalloc(&c->sort.traw, (size_t)(tiles + 1) * 4));
    HIPTRY(c, hipMalloc(&c->sort.flag, (size_t)nb * 4));
    HIPTRY(c, hipMemset(c->sort.flag, 0, (size_t)nb * 4));
    HIPTRY(c, hipMalloc(&c->sort.active, (size_t)nb * 4));
    HIPTRY(c, hipMemset(c->sort.active, 0, (size_t)nb * 4));
    HIPTRY(c, hipMalloc(&c->sort.bslot, (size_t)nb * 4));
    HIPTRY(c, hipMemset(c->sort.bslot, 0xff, (size_t)nb * 4));
    c->sort_nb_cap = nb;
    invalidate_graphs(c);
  }
  if (!c->sort.counters) {
    HIPTRY(c, hipMalloc(&c->sort.counters, 64));
    HIPTRY(c, hipMemset(c->sort.counters, 0, 64));
    const size_t rb = (size_t)kRunLanes * kRunStride * 4;
    HIPTRY(c, hipMalloc(&c->sort.runs, rb));
    HIPTRY(c, hipMemset(c->sort.runs, 0, rb));
    HIPTRY(c, hipHostGetDevicePointer((void**)&c->sort.runs_host, c->runs_h, 0));
    invalidate_graphs(c);
  }
  if (n > c->sort_n_cap || !c->sort.key) {
    for (void* p : {(void*)c->sort.key, (void*)c->sort.rank, (void*)c->sort.perm})
      if (p) hipFree(p);
    const size_t b = (size_t)(n ? n : 1) * 4;
    HIPTRY(c, hipMalloc(&c->sort.key, b));
    HIPTRY(c, hipMalloc(&c->sort.rank, b));
    HIPTRY(c, hipMalloc(&c->sort.perm, b));
    c->sort_n_cap = n;
    invalidate_graphs(c);
  }
  return AEMH_OK;
}

// Upload the scene table (contents only: graphs read it at replay time).
aemh_status set_scene_table(aemh_ctx* c, const std::vector<aemh_scene>& sc) {
  std::vector<SceneDev> d;
  d.reserve(sc.size());
  uint32_t off = 0, expect = 0;
  for (const auto& s : sc) {
    if (s.first != expect)
      return fail(c, AEMH_ERR_ARG, "scenes must tile the active set in order (first=%u, expected %u)",
                  s.first, expect);
    if (s.nx == 0 || s.ny == 0 || s.nz == 0 || !(s.dx > 0.0f))
      return fail(c, AEMH_ERR_ARG, "scene grid must be non-empty with dx > 0");
    expect += s.count;
    d.push_back(make_scene(s, c->cfg, off));
    uint64_t nbs = (uint64_t)d.back().nb[0] * d.back().nb[1] * d.back().nb[2];
    if (off + nbs >= (1ull << 26)) return fail(c, AEMH_ERR_ARG, "too many grid blocks");
    off += (uint32_t)nbs;
  }
  if (expect != c->n_work)
    return fail(c, AEMH_ERR_ARG, "scenes cover %u particles, active set has %u", expect, c->n_work);
  const bool shape_changed = d.size() != c->scd_h.size() || off != c->nb_total;
  c->nb_total = off;
  if (d.size() > c->scd_cap) {
    if (c->scd_d) hipFree(c->scd_d);
    c->scd_d = nullptr;
    c->scd_valid = false;
    HIPTRY(c, hipMalloc(&c->scd_d, d.size() * sizeof(SceneDev)));
    c->scd_cap = d.size();
    invalidate_graphs(c);
  }
  if (shape_changed) invalidate_graphs(c);
  // A repeated step with the same scene reuses the table already on device.
  const bool same = !shape_changed && c->scd_valid &&
                    memcmp(c->scd_h.data(), d.data(), d.size() * sizeof(SceneDev)) == 0;
  c->scd_h = d;
  if (!same) {
    drop_counts(c);  // keys are functions of the scene table
    HIPTRY(c, hipMemcpyAsync(c->scd_d, d.data(), d.size() * sizeof(SceneDev), hipMemcpyHostToDevice,
                             c->stream));
    c->scd_valid = true;
  }
  return ensure_grid_sort(c);
}

hipEvent_t get_event(aemh_ctx* c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  // No system-scope fence: a fenced record also writes back L2, which is
  // not needed for a timestamp on this stream.
  hipEvent_t e = nullptr;
  hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
  return e;
}

template <typename Fn>
void timed(aemh_ctx* c, const char* name, Fn&& fn) {
  if (!c->prof) {
    fn();
    return;
  }
  hipEvent_t a = get_event(c), b = get_event(c);
  (void)hipEventRecord(a, c->stream);
  fn();
  (void)hipEventRecord(b, c->stream);
  c->pending.push_back({name, {a, b}});
}

// timed() for a stage whose kernels launch through launch_k: the events ride
// on the kernels' own dispatches.  A stage that launched nothing (no
// particles) records an empty pair.
template <class Fn>
void timed_k(aemh_ctx* c, const char* name, Fn&& fn) {
  if (!c->prof) {
    fn();
    return;
  }
  hipEvent_t a = get_event(c), b = get_event(c);
  g_stage = StageEvents{a, b, false};
  fn();
  const bool launched = g_stage.launched;
  g_stage = StageEvents{};
  if (!launched) {
    (void)hipEventRecord(a, c->stream);
    (void)hipEventRecord(b, c->stream);
  }
  c->pending.push_back({name, {a, b}});
}

void collect_profile(aemh_ctx* c) {
  for (auto& pr : c->pending) {
    float ms = 0;
    if (hipEventElapsedTime(&ms, pr.second.first, pr.second.second) == hipSuccess) {
      auto& r = c->prof_rec[pr.first];
      r.total_ms += ms;
      r.count += 1;
    }
    c->ev_pool.push_back(pr.second.first);
    c->ev_pool.push_back(pr.second.second);
  }
  c->pending.clear();
}

// Average time of `reps` back-to-back runs of fn on the context stream.
template <typename Fn>
aemh_status time_launches(aemh_ctx* c, int reps, float* ms, Fn&& fn) {
  hipEvent_t a = get_event(c), b = get_event(c);
  fn();  // untimed: the first launch of a candidate pays cold caches and code loads
  HIPTRY(c, hipEventRecord(a, c->stream));
  for (int r = 0; r < reps; r++) fn();
  HIPTRY(c, hipEventRecord(b, c->stream));
  HIPTRY(c, hipEventSynchronize(b));
  HIPTRY(c, hipEventElapsedTime(ms, a, b));
  c->ev_pool.push_back(a);
  c->ev_pool.push_back(b);
  *ms /= (float)reps;
  return AEMH_OK;
}

aemh_status flush_work(aemh_ctx* c);

// Device index list of the work set with room for n entries.
aemh_status ensure_idx(aemh_ctx* c, uint32_t n) {
  const size_t need = n ? n : 1;
  if (c->idx_d && c->idx_cap >= need) return AEMH_OK;
  if (c->idx_d) hipFree(c->idx_d);
  c->idx_d = nullptr;
  c->idx_cap = 0;
  HIPTRY(c, hipMalloc(&c->idx_d, need * 4));
  c->idx_cap = need;
  return AEMH_OK;
}

// Device copy of a slot list (or permutation) for the lifecycle kernels.
aemh_status upload_slots(aemh_ctx* c, const uint32_t* slots, size_t n) {
  if (n > c->slots_cap) {
    if (c->slots_d) hipFree(c->slots_d);
    c->slots_d = nullptr;
    c->slots_cap = 0;
    HIPTRY(c, hipMalloc(&c->slots_d, n * 4));
    c->slots_cap = n;
  }
  if (n) HIPTRY(c, hipMemcpyAsync(c->slots_d, slots, n * 4, hipMemcpyHostToDevice, c->stream));
  return AEMH_OK;
}

aemh_status alloc_life(aemh_ctx* c, MpssDev& L, uint32_t cap) {
  for (void* p : {(void*)L.lifetime, (void*)L.age, (void*)L.active})
    if (p) hipFree(p);
  L = MpssDev{};
  HIPTRY(c, hipMalloc(&L.lifetime, (size_t)cap * 4));
  HIPTRY(c, hipMalloc(&L.age, (size_t)cap * 4));
  HIPTRY(c, hipMalloc(&L.active, (size_t)cap));
  return AEMH_OK;
}

// MpssBuffer::kill for each slot in order (mpss.rs:205-212): an active slot
// becomes inactive and goes on the free list.
aemh_status mpss_kill(aemh_ctx* c, const uint32_t* slots, uint32_t n) {
  std::vector<uint32_t> killed;
  for (uint32_t i = 0; i < n; i++) {
    const uint32_t k = slots[i];
    if (k < c->canon.cap && c->act_h[k]) {
      c->act_h[k] = 0;
      c->free_list.push_back(k);
      c->mpss_count -= 1;
      killed.push_back(k);
    }
  }
  if (killed.empty()) return AEMH_OK;
  aemh_status s = upload_slots(c, killed.data(), killed.size());
  if (s != AEMH_OK) return s;
  launch_mpss_set_active(c->stream, c->life.active, c->slots_d, (uint32_t)killed.size(), 0);
  CHECK_LAUNCH(c);
  HIPTRY(c, hipStreamSynchronize(c->stream));
  // The default work set (active slots below count) changed; an explicit
  // index list stays as the caller gave it.
  if (!c->idx_explicit) {
    if ((s = flush_work(c)) != AEMH_OK) return s;
    c->work_valid = false;
  }
  return AEMH_OK;
}

// Before the canonical slots are rewritten: the work buffers go back first
// and are gathered again at the next step.
aemh_status mpss_begin_write(aemh_ctx* c) {
  aemh_status s = flush_work(c);
  if (s != AEMH_OK) return s;
  c->work_valid = false;
  return AEMH_OK;
}

aemh_status mpss_ready(aemh_ctx* c) {
  if (!c) return AEMH_ERR_ARG;
  if (!c->mpss) return fail(c, AEMH_ERR_STATE, "no resident MpssBuffer (aemh_mpss_reserve first)");
  hipSetDevice(c->device);
  return AEMH_OK;
}

// logf of the host C library for every float in [0.1, 10] (kLnTab*), built
// once per process: the values the reference's f32::ln gets from the
// platform libm (constitutive.rs:384).
const std::vector<float>& host_logf_table() {
  static std::once_flag once;
  static std::vector<float> tab;
  std::call_once(once, [] {
    tab.resize(kLnTabSize);
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> ts;
    for (unsigned t = 0; t < nt; t++)
      ts.emplace_back([t, nt] {
        for (uint32_t i = t; i < kLnTabSize; i += nt) {
          const uint32_t b = kLnTabLo + i;
          float x;
          memcpy(&x, &b, 4);
          tab[i] = logf(x);
        }
      });
    for (auto& th : ts) th.join();
  });
  return tab;
}

int bit_length(uint64_t v) {
  int b = 0;
  while (v) {
    b++;
    v >>= 1;
  }
  return b;
}

// Buffers of the reference-order FP32 P2G for the current working set and
// grid: 27 entries per particle (keys and values double-buffered for the
// sort, 8 contribution floats each, 56 B per entry), the radix sort's
// temporary storage, and the logf table (uploaded once per context).
aemh_status ensure_serial(aemh_ctx* c) {
  SerialDev& sd = c->ser;
  if (!c->lntab_d) {
    const std::vector<float>& tab = host_logf_table();
    HIPTRY(c, hipMalloc(&c->lntab_d, (size_t)kLnTabSize * 4));
    HIPTRY(c, hipMemcpy(c->lntab_d, tab.data(), (size_t)kLnTabSize * 4, hipMemcpyHostToDevice));
  }
  sd.lntab = c->lntab_d;
  const size_t m = (size_t)(c->n_work ? c->n_work : 1) * 27;
  if (m >= (1ull << 31)) return fail(c, AEMH_ERR_ARG, "serial-order P2G: too many particles");
  sd.id_bits = std::max(1, bit_length(c->n_work ? c->n_work - 1 : 0));
  sd.end_bit = sd.id_bits + bit_length(c->grid.nodes);
  if (m > c->ser_cap) {
    for (void* p : {(void*)sd.key[0], (void*)sd.key[1], (void*)sd.val[0], (void*)sd.val[1],
                    (void*)sd.contrib, sd.tmp})
      if (p) hipFree(p);
    sd.key[0] = sd.key[1] = nullptr;
    sd.val[0] = sd.val[1] = nullptr;
    sd.contrib = nullptr;
    sd.tmp = nullptr;
    sd.tmp_bytes = 0;
    c->ser_cap = 0;
    for (int k = 0; k < 2; k++) {
      HIPTRY(c, hipMalloc(&sd.key[k], m * 8));
      HIPTRY(c, hipMalloc(&sd.val[k], m * 4));
    }
    HIPTRY(c, hipMalloc(&sd.contrib, m * 32));
    c->ser_cap = m;
  }
  // the temporary size is largest with all 64 key bits
  const size_t tb = sort_pairs_u64_temp_bytes((uint32_t)c->ser_cap, 64);
  if (tb > sd.tmp_bytes) {
    if (sd.tmp) hipFree(sd.tmp);
    sd.tmp = nullptr;
    sd.tmp_bytes = 0;
    HIPTRY(c, hipMalloc(&sd.tmp, tb));
    sd.tmp_bytes = tb;
  }
  return AEMH_OK;
}

// Launch one substep on the context stream (eager or under capture).
// One substep on work[cur].  A re-sort step writes G2P's result to
// work[cur ^ 1] in block order; any other step updates work[cur] in place.
// G2P of a cache-resident working set counts the next substep's sort; the
// count kernel runs when the previous substep's G2P did not (count_kernel).
void enqueue_substep(aemh_ctx* c, int cur, bool resort, bool count_kernel) {
  const PBuf a = c->work[cur].pb, b = resort ? c->work[cur ^ 1].pb : a;
  const int inplace = resort ? 0 : 1;
  const uint32_t n = c->n_work;
  const int ns = (int)c->scd_h.size();
  const int storage = c->work[0].storage;
  const int q32 = c->cfg.accum == AEMH_ACCUM_Q32_32;
  const int fuse = cache_resident(c) ? 1 : 0;
  hipStream_t st = c->stream;
  if (count_kernel)
    timed_k(c, "count", [&] { launch_count(st, a, c->scd_d, ns, n, c->sort, c->nb_total); });
  timed_k(c, "scan", [&] { launch_scan(st, c->scd_d, ns, c->sort, c->nb_total); });
  timed_k(c, "scatter", [&] { launch_scatter(st, n, c->nb_total, c->sort, c->grid, acc_layout(c)); });
  if (c->cfg.accum == AEMH_ACCUM_F32_SERIAL) {
    // FP32 sums in list order, the FP32 grid update (atomic-flush layout:
    // no P2G partials), the reference-order G2P of the deterministic build
    LaunchShape ls = c->shape;
    ls.p2g_partials = 0;
    timed_k(c, "p2g", [&] {
      det::launch_p2g_serial(st, a, c->scd_d, ns, c->mats_d, (int)c->nmat, n, c->ser, c->grid, storage);
    });
    timed_k(c, "grid_update", [&] { fast::launch_grid_update_mpf(st, ls, c->scd_d, ns, c->sort, c->grid, n, c->nb_total); });
    timed_k(c, "g2p", [&] {
      det::launch_g2p(st, ls, a, b, c->scd_d, ns, c->mats_d, (int)c->nmat, n, c->sort, c->grid, storage, inplace, fuse);
    });
  } else if (q32) {
    timed_k(c, "p2g", [&] {
      det::launch_p2g(st, c->shape, a, c->scd_d, ns, c->mats_d, (int)c->nmat, n, c->sort, c->grid, storage);
    });
    timed_k(c, "grid_update", [&] { det::launch_grid_update(st, c->shape, c->scd_d, ns, c->sort, c->grid, n, c->nb_total); });
    timed_k(c, "g2p", [&] {
      det::launch_g2p(st, c->shape, a, b, c->scd_d, ns, c->mats_d, (int)c->nmat, n, c->sort, c->grid, storage, inplace, fuse);
    });
  } else {
    timed_k(c, "p2g", [&] {
      fast::launch_p2g(st, c->shape, a, c->scd_d, ns, c->mats_d, (int)c->nmat, n, c->sort, c->grid, storage);
    });
    timed_k(c, "grid_update", [&] { fast::launch_grid_update(st, c->shape, c->scd_d, ns, c->sort, c->grid, n, c->nb_total); });
    timed_k(c, "g2p", [&] {
      fast::launch_g2p(st, c->shape, a, b, c->scd_d, ns, c->mats_d, (int)c->nmat, n, c->sort, c->grid, storage, inplace, fuse);
    });
  }
}

int graph_index(int cur, bool resort, bool count_kernel) {
  return cur * 4 + (resort ? 2 : 0) + (count_kernel ? 1 : 0);
}

aemh_status build_graph(aemh_ctx* c, int cur, bool resort, bool count_kernel) {
  hipGraph_t g = nullptr;
  HIPTRY(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  enqueue_substep(c, cur, resort, count_kernel);
  hipError_t e = hipStreamEndCapture(c->stream, &g);
  if (e != hipSuccess) return fail(c, AEMH_ERR_HIP, "graph capture: %s", hipGetErrorName(e));
  e = hipGraphInstantiate(&c->gexec[graph_index(cur, resort, count_kernel)], g, nullptr, nullptr, 0);
  hipGraphDestroy(g);
  if (e != hipSuccess) return fail(c, AEMH_ERR_HIP, "graph instantiate: %s", hipGetErrorName(e));
  c->graph_valid = true;
  return AEMH_OK;
}

aemh_status check_overflow(aemh_ctx* c) {
  if (c->cfg.accum != AEMH_ACCUM_Q32_32) return AEMH_OK;
  // The grid update flags any node whose sum of |contribution| reached 2^63
  // raw: only then could an int64 partial sum have wrapped in some order.
  HIPTRY(c, hipMemcpyAsync(c->flag_h, c->sort.counters + kCntOverflow, 4, hipMemcpyDeviceToHost,
                           c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  if (*c->flag_h) {
    HIPTRY(c, hipMemsetAsync(c->sort.counters + kCntOverflow, 0, 4, c->stream));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    return fail(c, AEMH_ERR_OVERFLOW,
                "Q32.32 accumulation out of range: a node's sum of |contribution| reached 2^31 "
                "(FixedPoint would saturate in an order-dependent way)");
  }
  return AEMH_OK;
}

// Reads the wave runs the last count left in host-mapped memory (the stream
// is synchronized) and decides whether the next substep re-sorts.
void update_sort_policy(aemh_ctx* c, bool counted_sorted) {
  const uint32_t runs = *(volatile uint32_t*)c->runs_h;
  if (counted_sorted)
    c->runs_base = runs;
  else if (c->runs_base && runs > c->runs_base + c->runs_base / 4)
    c->need_sort = true;
}

// Enqueues k substeps; with sync, waits for them and reports the Q32.32
// range check.  Without (aemh_step_async), the run-count re-sort trigger
// waits for the next synchronized call and the range check for aemh_sync
// or the next synchronous step (the overflow flag is sticky).
aemh_status run_substeps(aemh_ctx* c, uint32_t k, bool allow_graph = true, bool sync = true) {
  if (!c->mats_d) return fail(c, AEMH_ERR_STATE, "materials not set");
  if (c->n_work == 0) return AEMH_OK;
  const bool serial = c->cfg.accum == AEMH_ACCUM_F32_SERIAL;
  if (c->shape.p2g_partials && !serial) {
    aemh_status s = ensure_partials(c);
    if (s != AEMH_OK) return s;
  }
  if (serial) {
    aemh_status s = ensure_serial(c);
    if (s != AEMH_OK) return s;
  }
  // Steps replay the captured substep graph.  With profiling on they launch
  // eagerly with an event pair around each kernel section: event-record
  // nodes inside a graph cost the replay more than eager launches do.  The
  // reference-order mode launches eagerly: its library radix sort is not
  // captured.
  const bool graph = allow_graph && c->use_graph && !c->prof && !serial;
  bool counted_sorted = false;  // the last substep's count saw a re-sorted set
  for (uint32_t i = 0; i < k; i++) {
    // Re-sort on the first substep of a new working set, when the block
    // order has decayed, and at least every kResortMax substeps.
    const uint32_t every = c->cfg.resort_interval;
    const bool resident = cache_resident(c);
    const bool resort = every ? (c->need_sort || c->since_sort + 1 >= every)
                              : (c->need_sort ||
                                 c->since_sort + 1 >= (resident ? kResortResident : kResortMax));
    counted_sorted = c->fresh_sort;
    const bool count_kernel = !c->counted;
    if (graph) {
      hipGraphExec_t& ge = c->gexec[graph_index(c->cur, resort, count_kernel)];
      if (!ge) {
        aemh_status s = build_graph(c, c->cur, resort, count_kernel);
        if (s != AEMH_OK) return s;
      }
      HIPTRY(c, hipGraphLaunch(ge, c->stream));
    } else {
      enqueue_substep(c, c->cur, resort, count_kernel);
      CHECK_LAUNCH(c);
    }
    c->counted = resident;  // G2P counted the next substep's sort
    if (resort) {
      c->cur ^= 1;
      c->since_sort = 0;
      c->need_sort = false;
    } else {
      c->since_sort += 1;
    }
    c->fresh_sort = resort;
  }
  c->work_dirty = true;
  if (!sync && !c->prof) {
    c->unsynced = true;
    return AEMH_OK;
  }
  HIPTRY(c, hipStreamSynchronize(c->stream));
  c->unsynced = false;
  update_sort_policy(c, counted_sorted);
  if (c->prof) collect_profile(c);
  return check_overflow(c);
}

aemh_status ensure_work(aemh_ctx* c) {
  if (c->work_valid && (c->idx_explicit || c->work_default || !c->mpss)) return AEMH_OK;
  if (!c->canon.slab) return fail(c, AEMH_ERR_STATE, "no particles uploaded");
  if (c->idx_explicit) {
    const std::vector<uint32_t> keep = c->idx_h;
    return aemh_set_indices(c, keep.data(), (uint32_t)keep.size());
  }
  // No list given: every slot; for a resident MpssBuffer the slots the serial
  // substep visits, i < count and active, ascending (mpm_solver.rs:187-193).
  std::vector<uint32_t> all;
  all.reserve(c->canon.cap);
  const uint32_t lim = c->mpss ? c->mpss_count : c->canon.cap;
  for (uint32_t i = 0; i < lim; i++)
    if (!c->mpss || c->act_h[i]) all.push_back(i);
  aemh_status s = aemh_set_indices(c, all.data(), (uint32_t)all.size());
  c->idx_explicit = false;
  c->work_default = s == AEMH_OK;
  return s;
}

aemh_status flush_work(aemh_ctx* c) {
  if (!c->work_valid || !c->work_dirty) return AEMH_OK;
  launch_scatter_work(c->stream, c->work[c->cur].pb, c->canon.pb, c->idx_d, c->n_work,
                      c->work[0].storage);
  CHECK_LAUNCH(c);
  HIPTRY(c, hipStreamSynchronize(c->stream));
  c->work_dirty = false;
  return AEMH_OK;
}

// Scene table of a plain (single-grid) step.
aemh_status prepare_single_scene(aemh_ctx* c, float sub_dt, int vel_only) {
  if (!c->grid_set) return fail(c, AEMH_ERR_STATE, "grid not set");
  aemh_status s = ensure_work(c);
  if (s != AEMH_OK) return s;
  aemh_scene sc{};
  sc.first = 0;
  sc.count = c->n_work;
  sc.nx = c->nx;
  sc.ny = c->ny;
  sc.nz = c->nz;
  sc.dx = c->dx;
  for (int a = 0; a < 3; a++) sc.origin[a] = c->origin[a];
  sc.dt = sub_dt;
  sc.velocity_only = vel_only;
  sc.clamp_mode = c->cfg.clamp_mode;
  c->batch_mode = false;
  return set_scene_table(c, {sc});
}

aemh_status single_scene(aemh_ctx* c, float total_dt, int vel_only, bool sync = true) {
  if (c->cfg.substeps == 0) return fail(c, AEMH_ERR_ARG, "substeps must be >= 1");
  if (!c->grid_set) return fail(c, AEMH_ERR_STATE, "grid not set");
  aemh_status s = ensure_work(c);
  if (s != AEMH_OK) return s;
  if (c->n_work == 0) return AEMH_OK;  // nothing to advance
  s = prepare_single_scene(c, total_dt / (float)c->cfg.substeps,  // mpm_solver.rs:1075
                                       vel_only);
  if (s != AEMH_OK) return s;
  return run_substeps(c, c->cfg.substeps, true, sync);
}

}  // namespace

namespace {
__global__ void k_aos_to_soa_v(const float* __restrict__ aos, Ptr9 soa, int comps, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < comps; c++) soa.p[c][i] = aos[(size_t)i * comps + c];
}
__global__ void k_soa_to_aos_v(Ptr9 soa, float* __restrict__ aos, int comps, uint32_t n) {
  const uint32_t i = blockIdx.x * 256u + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < comps; c++) aos[(size_t)i * comps + c] = soa.p[c][i];
}

aemh_status upload_aos(aemh_ctx* c, const float* host, float* const* dst, int comps, uint32_t n) {
  if (!host) return AEMH_OK;
  aemh_status s = ensure_staging(c, (size_t)n * comps * 4);
  if (s != AEMH_OK) return s;
  HIPTRY(c, hipMemcpyAsync(c->staging, host, (size_t)n * comps * 4, hipMemcpyHostToDevice, c->stream));
  Ptr9 p{};
  for (int k = 0; k < comps; k++) p.p[k] = dst[k];
  k_aos_to_soa_v<<<(n + 255) / 256, 256, 0, c->stream>>>(c->staging, p, comps, n);
  CHECK_LAUNCH(c);
  return AEMH_OK;
}
aemh_status download_aos(aemh_ctx* c, float* host, float* const* src, int comps, uint32_t n) {
  if (!host) return AEMH_OK;
  aemh_status s = ensure_staging(c, (size_t)n * comps * 4);
  if (s != AEMH_OK) return s;
  Ptr9 p{};
  for (int k = 0; k < comps; k++) p.p[k] = src[k];
  k_soa_to_aos_v<<<(n + 255) / 256, 256, 0, c->stream>>>(p, c->staging, comps, n);
  CHECK_LAUNCH(c);
  HIPTRY(c, hipMemcpyAsync(host, c->staging, (size_t)n * comps * 4, hipMemcpyDeviceToHost, c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  return AEMH_OK;
}


// ---------------------------------------------------------------------------
// LOD tier orchestrator helpers (simulation.rs:480-700)
// ---------------------------------------------------------------------------
aemh_status ensure_tier(aemh_ctx* c) {
  const uint32_t cap = c->canon.cap;
  if (c->tier_cap == cap && c->tier_cls) return AEMH_OK;
  for (void* p : {(void*)c->tier_list[0], (void*)c->tier_list[1], (void*)c->tier_list[2],
                  (void*)c->tier_cls, (void*)c->tier_dist, (void*)c->tier_keys[0],
                  (void*)c->tier_keys[1], (void*)c->tier_vals, (void*)c->tier_bcnt,
                  c->tier_sort_tmp, (void*)c->ser.key[0], (void*)c->ser.key[1],
                  (void*)c->ser.val[0], (void*)c->ser.val[1], (void*)c->ser.contrib, c->ser.tmp,
                  (void*)c->lntab_d})
    if (p) hipFree(p);
  c->tier_cap = 0;
  for (int t = 0; t < 3; t++) c->tier_n[t] = 0;
  const size_t n = cap ? cap : 1;
  for (int t = 0; t < 3; t++) HIPTRY(c, hipMalloc(&c->tier_list[t], n * 4));
  HIPTRY(c, hipMalloc(&c->tier_cls, n));
  HIPTRY(c, hipMalloc(&c->tier_dist, n * 4));
  HIPTRY(c, hipMalloc(&c->tier_keys[0], n * 4));
  HIPTRY(c, hipMalloc(&c->tier_keys[1], n * 4));
  HIPTRY(c, hipMalloc(&c->tier_vals, n * 4));
  HIPTRY(c, hipMalloc(&c->tier_bcnt, ((size_t)select_tiles((uint32_t)n) * 6 + 4) * 4));
  c->tier_sort_bytes = sort_pairs_temp_bytes((uint32_t)n);
  HIPTRY(c, hipMalloc(&c->tier_sort_tmp, c->tier_sort_bytes ? c->tier_sort_bytes : 4));
  c->tier_cap = cap;
  return AEMH_OK;
}

// Ordered compaction through the tile counters; *tot3 (host) gets the sizes.
aemh_status tier_select(aemh_ctx* c, const uint8_t* cls, const uint32_t* src, const float* key,
                        uint32_t n, uint32_t* out0, uint32_t* out1, uint32_t* out2,
                        uint32_t* key0, uint32_t tot3[3]) {
  const size_t tiles = select_tiles(n ? n : 1);
  uint32_t* bcnt = c->tier_bcnt;
  uint32_t* boff = bcnt + tiles * 3;
  uint32_t* tot = boff + tiles * 3;
  timed(c, "lod_select", [&] {
    launch_select(c->stream, cls, src, key, n, bcnt, boff, tot, out0, out1, out2, key0);
  });
  CHECK_LAUNCH(c);
  HIPTRY(c, hipMemcpyAsync(tot3, tot, 12, hipMemcpyDeviceToHost, c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  return AEMH_OK;
}

// LOD reclassification (simulation.rs:495-531): near / mid / far lists of the
// active slots below count in ascending slot order; a near list longer than
// max_near_particles keeps its max_near closest (ties: lower slot first, as
// a stable sort by distance gives) and the rest join the mid list.
aemh_status tier_classify(aemh_ctx* c, const float player[3]) {
  const aemh_lod_config& L = c->lod;
  const float near2 = L.near_field_distance * L.near_field_distance;
  const float mid2 = L.mid_field_distance * L.mid_field_distance;
  const uint32_t count = c->mpss_count;
  aemh_status s = flush_work(c);  // positions as the last step left them
  if (s != AEMH_OK) return s;
  timed(c, "lod_classify", [&] {
    launch_lod_classify(c->stream, c->canon.pb, c->life.active, count, player, near2, mid2,
                        c->tier_cls, c->tier_dist);
  });
  CHECK_LAUNCH(c);
  uint32_t n3[3] = {0, 0, 0};
  s = tier_select(c, c->tier_cls, nullptr, c->tier_dist, count, c->tier_list[0],
                              c->tier_list[1], c->tier_list[2], c->tier_keys[0], n3);
  if (s != AEMH_OK) return s;
  const uint32_t max_near = L.max_near_particles;
  if (n3[0] > max_near) {
    hipError_t e = hipSuccess;
    timed(c, "near_cull", [&] {
      e = launch_sort_pairs(c->stream, c->tier_sort_tmp, c->tier_sort_bytes, c->tier_keys[0],
                            c->tier_keys[1], c->tier_list[0], c->tier_vals, n3[0]);
    });
    HIPTRY(c, e);
    if (max_near)
      HIPTRY(c, hipMemcpyAsync(c->tier_list[0], c->tier_vals, (size_t)max_near * 4,
                               hipMemcpyDeviceToDevice, c->stream));
    HIPTRY(c, hipMemcpyAsync(c->tier_list[1] + n3[1], c->tier_vals + max_near,
                             (size_t)(n3[0] - max_near) * 4, hipMemcpyDeviceToDevice, c->stream));
    HIPTRY(c, hipStreamSynchronize(c->stream));
    n3[1] += n3[0] - max_near;
    n3[0] = max_near;
  }
  for (int t = 0; t < 3; t++) c->tier_n[t] = n3[t];
  c->tier_version += 1;
  return AEMH_OK;
}

// Work set = the given device lists back to back (gathered from the
// canonical slots); reused while the same lists are asked for again.
aemh_status tier_set_work(aemh_ctx* c, uint64_t key, const uint32_t* const* lists,
                          const uint32_t* counts, int nl) {
  if (c->work_valid && c->work_tier_key == key) return AEMH_OK;
  aemh_status s = flush_work(c);
  if (s != AEMH_OK) return s;
  uint32_t n = 0;
  for (int k = 0; k < nl; k++) n += counts[k];
  if (n >= kMaxWork) return fail(c, AEMH_ERR_ARG, "working set of %u particles: limit 2^30 - 1", n);
  const int storage = (int)c->cfg.storage;
  for (int k = 0; k < 2; k++)
    if ((s = alloc_store(c, c->work[k], n ? n : 1, storage)) != AEMH_OK) return s;
  if ((s = ensure_idx(c, n)) != AEMH_OK) return s;
  uint32_t off = 0;
  for (int k = 0; k < nl; k++) {
    if (counts[k])
      HIPTRY(c, hipMemcpyAsync(c->idx_d + off, lists[k], (size_t)counts[k] * 4,
                               hipMemcpyDeviceToDevice, c->stream));
    off += counts[k];
  }
  launch_gather_work(c->stream, c->canon.pb, c->work[0].pb, c->idx_d, n, storage);
  CHECK_LAUNCH(c);
  HIPTRY(c, hipStreamSynchronize(c->stream));
  c->n_work = n;
  c->cur = 0;
  c->need_sort = true;  // list order: the first substep re-sorts
  c->runs_base = 0;
  c->idx_h.clear();
  c->idx_explicit = false;
  c->work_default = false;
  c->work_valid = true;
  c->work_dirty = false;
  c->batch_mode = false;
  c->work_tier_key = key;
  invalidate_graphs(c);
  return AEMH_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

void aemh_config_default(aemh_config* cfg) {
  if (!cfg) return;
  memset(cfg, 0, sizeof(*cfg));
  // MpmConfig::default (mpm_solver.rs:57-69)
  cfg->substeps = 1;
  cfg->gravity[0] = 0.0f;
  cfg->gravity[1] = -9.81f;
  cfg->gravity[2] = 0.0f;
  cfg->damping = 0.995f;
  cfg->bc_margin = 0;
  cfg->bc_restitution = 0.0f;
  cfg->accum = AEMH_ACCUM_F32;
  cfg->clamp_mode = AEMH_CLAMP_AUTO;
  cfg->storage = AEMH_STORE_F32;
}

const char* aemh_version(void) { return kVersion; }

aemh_status aemh_create(aemh_ctx** out, int device) {
  if (!out) return AEMH_ERR_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return AEMH_ERR_NO_DEVICE;
  if (device < 0 || device >= count) return AEMH_ERR_ARG;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return AEMH_ERR_NO_DEVICE;
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return AEMH_ERR_NO_DEVICE;
  aemh_ctx* c = new aemh_ctx();
  c->device = device;
  aemh_config_default(&c->cfg);
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
      hipHostMalloc((void**)&c->flag_h, 8, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->runs_h, 64, hipHostMallocMapped | hipHostMallocCoherent) !=
          hipSuccess) {
    delete c;
    return AEMH_ERR_HIP;
  }
  *c->runs_h = 0;
  // Only aemh_conservation_f64_noalloc needs it: a failure is kept and
  // reported by that entry point.
  const hipError_t se = scratch_for_device(device, &c->cons_scratch);
  if (se != hipSuccess) {
    c->cons_scratch = nullptr;
    c->scratch_err = hipGetErrorName(se);
  }
  *out = c;
  return AEMH_OK;
}

aemh_status aemh_destroy(aemh_ctx* c) {
  if (!c) return AEMH_OK;
  hipSetDevice(c->device);
  hipStreamSynchronize(c->stream);
  c->counted = false;  // nothing to clear: the buffers go
  invalidate_graphs(c);
  for (auto& pr : c->pending) {
    hipEventDestroy(pr.second.first);
    hipEventDestroy(pr.second.second);
  }
  for (auto e : c->ev_pool) hipEventDestroy(e);
  for (void* p : {(void*)c->mats_d, c->canon.slab, c->work[0].slab, c->work[1].slab,
                  (void*)c->idx_d, (void*)c->staging, (void*)c->scd_d, (void*)c->grid.accf,
                  (void*)c->grid.accq, (void*)c->grid.gv, (void*)c->sort.key, (void*)c->sort.rank,
                  (void*)c->sort.perm, (void*)c->sort.cnt, (void*)c->sort.off, (void*)c->sort.tsum, (void*)c->sort.traw,
                  (void*)c->sort.flag, (void*)c->sort.active, (void*)c->sort.counters,
                  (void*)c->sort.runs,
                  (void*)c->sort.bslot, (void*)c->grid.partf, (void*)c->grid.partq,
                  (void*)c->life.lifetime, (void*)c->life.age, (void*)c->life.active,
                  (void*)c->life_tmp.lifetime, (void*)c->life_tmp.age, (void*)c->life_tmp.active,
                  c->canon_tmp.slab, (void*)c->slots_d, (void*)c->expired_d,
                  (void*)c->tier_list[0], (void*)c->tier_list[1], (void*)c->tier_list[2],
                  (void*)c->tier_cls, (void*)c->tier_dist, (void*)c->tier_keys[0],
                  (void*)c->tier_keys[1], (void*)c->tier_vals, (void*)c->tier_bcnt,
                  c->tier_sort_tmp})
    if (p) hipFree(p);
  if (c->flag_h) hipHostFree(c->flag_h);
  if (c->runs_h) hipHostFree(c->runs_h);
  hipStreamDestroy(c->stream);
  delete c;
  return AEMH_OK;
}

const char* aemh_last_error(const aemh_ctx* c) { return c ? c->err.c_str() : "null context"; }

aemh_status aemh_set_config(aemh_ctx* c, const aemh_config* cfg) {
  if (!c || !cfg) return fail(c, AEMH_ERR_ARG, "null argument");
  if (cfg->substeps == 0) return fail(c, AEMH_ERR_ARG, "substeps must be >= 1");
  if (cfg->accum > AEMH_ACCUM_F32_SERIAL || cfg->clamp_mode > AEMH_CLAMP_WINDOW ||
      cfg->storage > AEMH_STORE_F16)
    return fail(c, AEMH_ERR_ARG, "unknown mode value");
  hipSetDevice(c->device);
  const bool storage_changed = cfg->storage != c->cfg.storage;
  if (storage_changed && c->work_valid) {
    aemh_status s = flush_work(c);
    if (s != AEMH_OK) return s;
    c->work_valid = false;
  }
  if (cfg->accum != c->cfg.accum || storage_changed) invalidate_graphs(c);
  c->cfg = *cfg;
  return AEMH_OK;
}

aemh_status aemh_set_grid(aemh_ctx* c, uint32_t nx, uint32_t ny, uint32_t nz, float dx,
                          const float origin[3]) {
  if (!c) return AEMH_ERR_ARG;
  if (nx == 0 || ny == 0 || nz == 0 || !(dx > 0.0f) || !std::isfinite(dx))
    return fail(c, AEMH_ERR_ARG, "grid dims must be > 0 and dx > 0");
  if ((uint64_t)nx * ny * nz > (1ull << 31)) return fail(c, AEMH_ERR_ARG, "grid too large");
  c->nx = nx;
  c->ny = ny;
  c->nz = nz;
  c->dx = dx;
  for (int a = 0; a < 3; a++) c->origin[a] = origin ? origin[a] : 0.0f;
  c->grid_set = true;
  return AEMH_OK;
}

aemh_status aemh_set_materials(aemh_ctx* c, uint32_t n, const float* params7,
                               const uint8_t* models, const float* ext3) {
  if (!c) return AEMH_ERR_ARG;
  if (n == 0 || !params7 || !models)
    return fail(c, AEMH_ERR_ARG, "material table must be non-empty (material_table[0] is the fallback)");
  hipSetDevice(c->device);
  std::vector<MatDev> m(n + 1);
  for (uint32_t i = 0; i < n; i++)
    m[i] = make_mat(params7 + 7 * (size_t)i, models[i], ext3 ? ext3 + 3 * (size_t)i : nullptr);
  // unwrap_or(&material_table[0]) with NeoHookean (mpm_solver.rs:840-841)
  m[n] = make_mat(params7, AEMH_MODEL_NEO_HOOKEAN, ext3);
  if (n + 1 > c->mats_h.size() || !c->mats_d) {
    if (c->mats_d) hipFree(c->mats_d);
    c->mats_d = nullptr;
    HIPTRY(c, hipMalloc(&c->mats_d, (n + 1) * sizeof(MatDev)));
    invalidate_graphs(c);
  }
  if (n != c->nmat) invalidate_graphs(c);
  c->mats_h = m;
  c->nmat = n;
  HIPTRY(c, hipMemcpy(c->mats_d, m.data(), (n + 1) * sizeof(MatDev), hipMemcpyHostToDevice));
  return AEMH_OK;
}

aemh_status aemh_upload_particles(aemh_ctx* c, uint32_t n, const float* pos3, const float* vel3,
                                  const float* C9, const float* F9, const float* J,
                                  const float* mass, const uint16_t* mat) {
  if (!c) return AEMH_ERR_ARG;
  if (n == 0 || !pos3) return fail(c, AEMH_ERR_ARG, "need n > 0 and positions");
  hipSetDevice(c->device);
  aemh_status s = alloc_store(c, c->canon, n, 0);
  if (s != AEMH_OK) return s;
  c->canon.cap = n;
  PBuf& P = c->canon.pb;
  // MpssBuffer::new / reset_slot defaults (mpss.rs:161-185, 215-232)
  launch_fill(c->stream, P.J, 1.0f, n);
  launch_fill(c->stream, P.Jp, 1.0f, n);
  launch_fill(c->stream, P.mass, 1.0f, n);
  for (int a = 0; a < 3; a++) launch_fill(c->stream, P.v[a], 0.0f, n);
  for (int k = 0; k < 9; k++) {
    launch_fill(c->stream, P.C[k], 0.0f, n);
    launch_fill(c->stream, P.F[k], (k % 4 == 0) ? 1.0f : 0.0f, n);
  }
  HIPTRY(c, hipMemsetAsync(P.mat, 0, (size_t)n * 2, c->stream));
  if ((s = upload_aos(c, pos3, P.x, 3, n)) != AEMH_OK) return s;
  if ((s = upload_aos(c, vel3, P.v, 3, n)) != AEMH_OK) return s;
  if ((s = upload_aos(c, C9, P.C, 9, n)) != AEMH_OK) return s;
  if ((s = upload_aos(c, F9, P.F, 9, n)) != AEMH_OK) return s;
  if (J) HIPTRY(c, hipMemcpyAsync(P.J, J, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
  if (mass) HIPTRY(c, hipMemcpyAsync(P.mass, mass, (size_t)n * 4, hipMemcpyHostToDevice, c->stream));
  if (mat) HIPTRY(c, hipMemcpyAsync(P.mat, mat, (size_t)n * 2, hipMemcpyHostToDevice, c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  CHECK_LAUNCH(c);
  c->work_valid = false;
  c->work_dirty = false;
  c->idx_explicit = false;
  c->mpss = false;
  c->work_tier_key = 0;
  for (int t = 0; t < 3; t++) c->tier_n[t] = 0;
  if (c->sort.counters) HIPTRY(c, hipMemset(c->sort.counters, 0, 64));
  return AEMH_OK;
}

aemh_status aemh_upload_plastic_j(aemh_ctx* c, const float* Jp) {
  if (!c || !Jp) return fail(c, AEMH_ERR_ARG, "null argument");
  if (!c->canon.slab) return fail(c, AEMH_ERR_STATE, "no particles uploaded");
  hipSetDevice(c->device);
  aemh_status s = flush_work(c);
  if (s != AEMH_OK) return s;
  HIPTRY(c, hipMemcpy(c->canon.pb.Jp, Jp, (size_t)c->canon.cap * 4, hipMemcpyHostToDevice));
  c->work_valid = false;
  return AEMH_OK;
}

aemh_status aemh_download_particles(aemh_ctx* c, float* pos3, float* vel3, float* C9, float* F9,
                                    float* J) {
  if (!c) return AEMH_ERR_ARG;
  if (!c->canon.slab) return fail(c, AEMH_ERR_STATE, "no particles uploaded");
  hipSetDevice(c->device);
  aemh_status s = flush_work(c);
  if (s != AEMH_OK) return s;
  const uint32_t n = c->canon.cap;
  PBuf& P = c->canon.pb;
  if ((s = download_aos(c, pos3, P.x, 3, n)) != AEMH_OK) return s;
  if ((s = download_aos(c, vel3, P.v, 3, n)) != AEMH_OK) return s;
  if ((s = download_aos(c, C9, P.C, 9, n)) != AEMH_OK) return s;
  if ((s = download_aos(c, F9, P.F, 9, n)) != AEMH_OK) return s;
  if (J) HIPTRY(c, hipMemcpy(J, P.J, (size_t)n * 4, hipMemcpyDeviceToHost));
  return AEMH_OK;
}

aemh_status aemh_download_plastic_j(aemh_ctx* c, float* Jp) {
  if (!c || !Jp) return fail(c, AEMH_ERR_ARG, "null argument");
  if (!c->canon.slab) return fail(c, AEMH_ERR_STATE, "no particles uploaded");
  hipSetDevice(c->device);
  aemh_status s = flush_work(c);
  if (s != AEMH_OK) return s;
  HIPTRY(c, hipMemcpy(Jp, c->canon.pb.Jp, (size_t)c->canon.cap * 4, hipMemcpyDeviceToHost));
  return AEMH_OK;
}

aemh_status aemh_set_indices(aemh_ctx* c, const uint32_t* idx, uint32_t n_idx) {
  if (!c) return AEMH_ERR_ARG;
  if (!c->canon.slab) return fail(c, AEMH_ERR_STATE, "no particles uploaded");
  if (n_idx && !idx) return fail(c, AEMH_ERR_ARG, "null index list");
  if (n_idx >= kMaxWork) return fail(c, AEMH_ERR_ARG, "working set of %u particles: limit 2^30 - 1", n_idx);
  hipSetDevice(c->device);
  // Distinct, in-range indices (the reference relies on this by convention,
  // mpm_solver.rs:1462; a duplicate would make two G2P writers race here).
  std::vector<uint8_t> seen(c->canon.cap, 0);
  for (uint32_t i = 0; i < n_idx; i++) {
    if (idx[i] >= c->canon.cap)
      return fail(c, AEMH_ERR_ARG, "index %u out of range (capacity %u)", idx[i], c->canon.cap);
    if (seen[idx[i]]) return fail(c, AEMH_ERR_ARG, "duplicate particle index %u", idx[i]);
    seen[idx[i]] = 1;
  }
  aemh_status s = flush_work(c);
  if (s != AEMH_OK) return s;
  const int storage = (int)c->cfg.storage;
  for (int k = 0; k < 2; k++)
    if ((s = alloc_store(c, c->work[k], n_idx ? n_idx : 1, storage)) != AEMH_OK) return s;
  if ((s = ensure_idx(c, n_idx)) != AEMH_OK) return s;
  if (n_idx) HIPTRY(c, hipMemcpy(c->idx_d, idx, (size_t)n_idx * 4, hipMemcpyHostToDevice));
  c->work_tier_key = 0;
  c->idx_h.assign(idx, idx + n_idx);
  c->idx_explicit = true;
  c->work_default = false;
  c->n_work = n_idx;
  c->cur = 0;
  c->need_sort = true;  // list order: the first substep re-sorts
  c->runs_base = 0;
  launch_gather_work(c->stream, c->canon.pb, c->work[0].pb, c->idx_d, n_idx, storage);
  CHECK_LAUNCH(c);
  HIPTRY(c, hipStreamSynchronize(c->stream));
  c->work_valid = true;
  c->work_dirty = false;
  c->batch_mode = false;
  invalidate_graphs(c);
  return AEMH_OK;
}

aemh_status aemh_step(aemh_ctx* c, float total_dt) {
  if (!c) return AEMH_ERR_ARG;
  hipSetDevice(c->device);
  aemh_status s = single_scene(c, total_dt, 0);
  if (s != AEMH_OK || !c->mpss) return s;
  uint32_t n_exp = 0;  // mpm_step_with_grid ends with update_lifetimes (:508)
  return aemh_mpss_update_lifetimes(c, total_dt, nullptr, 0, &n_exp);
}

aemh_status aemh_step_async(aemh_ctx* c, float total_dt) {
  if (!c) return AEMH_ERR_ARG;
  // A resident MpssBuffer ages its lifetimes after the step, which reads the
  // expiry count back: that step stays synchronous.
  if (c->mpss) return aemh_step(c, total_dt);
  hipSetDevice(c->device);
  return single_scene(c, total_dt, 0, false);
}

aemh_status aemh_sync(aemh_ctx* c) {
  if (!c) return AEMH_ERR_ARG;
  hipSetDevice(c->device);
  HIPTRY(c, hipStreamSynchronize(c->stream));
  if (!c->unsynced) return AEMH_OK;
  c->unsynced = false;
  update_sort_policy(c, c->fresh_sort);
  return check_overflow(c);
}

aemh_status aemh_mpm_step_parallel_with_grid(aemh_ctx* c, const uint32_t* idx, uint32_t n_idx,
                                             float total_dt) {
  if (!c) return AEMH_ERR_ARG;
  if (n_idx == 0) return AEMH_OK;  // mpm_solver.rs:1072-1074
  if (!idx) return fail(c, AEMH_ERR_ARG, "null index list");
  hipSetDevice(c->device);
  if (!c->work_valid || c->idx_h.size() != n_idx ||
      memcmp(c->idx_h.data(), idx, (size_t)n_idx * 4) != 0) {
    aemh_status s = aemh_set_indices(c, idx, n_idx);
    if (s != AEMH_OK) return s;
  }
  c->idx_explicit = true;
  aemh_status s = single_scene(c, total_dt, 0);
  // With a resident buffer the list holds for this call only, as the
  // reference takes particle_indices per call.
  if (c->mpss) c->idx_explicit = false;
  if (s != AEMH_OK || !c->mpss) return s;
  // buffer.update_lifetimes(total_dt) closes the reference step (:1083).
  uint32_t n_exp = 0;
  return aemh_mpss_update_lifetimes(c, total_dt, nullptr, 0, &n_exp);
}

aemh_status aemh_step_velocity_only(aemh_ctx* c, float total_dt) {
  if (!c) return AEMH_ERR_ARG;
  hipSetDevice(c->device);
  return single_scene(c, total_dt, 1);
}

aemh_status aemh_set_scenes(aemh_ctx* c, const aemh_scene* scenes, uint32_t n) {
  if (!c) return AEMH_ERR_ARG;
  if (n == 0 || !scenes) return fail(c, AEMH_ERR_ARG, "need at least one scene");
  hipSetDevice(c->device);
  aemh_status s = ensure_work(c);
  if (s != AEMH_OK) return s;
  c->scenes_h.assign(scenes, scenes + n);
  s = set_scene_table(c, c->scenes_h);
  if (s != AEMH_OK) return s;
  c->batch_mode = true;
  return AEMH_OK;
}

namespace {
// Before any batched launch: calls made since aemh_set_scenes may have
// replaced the particles (upload, J_p upload, storage change: the working
// set is stale) or changed the configuration the scene table and the grid
// buffers were built from (accumulation mode, gravity, damping, walls).
// The working set is gathered again and the table rebuilt from the scenes
// as given; a table that no longer tiles the new working set is an error.
// An unchanged table is neither re-uploaded nor re-allocated.
aemh_status revalidate_batch(aemh_ctx* c) {
  if (!c->work_valid) {
    aemh_status s = ensure_work(c);
    if (s != AEMH_OK) return s;
  }
  aemh_status s = set_scene_table(c, c->scenes_h);
  if (s != AEMH_OK) return s;
  c->batch_mode = true;
  return AEMH_OK;
}
}  // namespace

aemh_status aemh_batch_step(aemh_ctx* c, uint32_t n_substeps) {
  if (!c) return AEMH_ERR_ARG;
  if (!c->batch_mode) return fail(c, AEMH_ERR_STATE, "aemh_set_scenes first");
  hipSetDevice(c->device);
  aemh_status r = revalidate_batch(c);
  if (r != AEMH_OK) return r;
  if (!c->tuned) {
    aemh_status s = aemh_autotune(c);
    if (s != AEMH_OK) return s;
  }
  return run_substeps(c, n_substeps);
}

// Times the candidate launch shapes of P2G (LDS tiles per workgroup and tile
// flush mode, with the grid update that consumes it), the grid update
// (workgroup count) and G2P (workgroup size) on the current particles and
// scenes and keeps the fastest of each.  The substep kernels
// run without advancing the state: G2P, timed in its in-place form, writes
// the idle ping-pong buffer (which the next re-sorting step overwrites) and
// the sort and grid buffers are rebuilt by the next step.
aemh_status aemh_autotune(aemh_ctx* c) {
  if (!c) return AEMH_ERR_ARG;
  if (!c->mats_d) return fail(c, AEMH_ERR_STATE, "materials not set");
  hipSetDevice(c->device);
  aemh_status s = c->batch_mode ? revalidate_batch(c)
                                : prepare_single_scene(c, c->scd_h.empty() ? 1e-4f : c->scd_h[0].dt, 0);
  if (s != AEMH_OK) return s;
  if (c->n_work == 0) return AEMH_OK;
  const bool prof = c->prof;
  c->prof = false;
  const PBuf a = c->work[c->cur].pb, b = c->work[c->cur ^ 1].pb;
  const uint32_t n = c->n_work;
  const int ns = (int)c->scd_h.size();
  const int storage = c->work[0].storage;
  const int q32 = c->cfg.accum == AEMH_ACCUM_Q32_32;
  hipStream_t st = c->stream;
  LaunchShape& ls = c->shape;
  drop_counts(c);  // the timing runs count and sort afresh
  launch_count(st, a, c->scd_d, ns, n, c->sort, c->nb_total);
  launch_scan(st, c->scd_d, ns, c->sort, c->nb_total);
  launch_scatter(st, n, c->nb_total, c->sort, c->grid, acc_layout(c));
  CHECK_LAUNCH(c);
  constexpr int kReps = 3;
  float best = 1e30f, best_ms = 0.0f, ms = 0.0f;
  int pick = ls.p2g_tiles, pick_part = ls.p2g_partials, pick_pairs = ls.p2g_pairs,
      pick_layers = ls.p2g_layers;
  // Tile budget, flush mode and lane assignment together: the flush mode
  // moves work between P2G and the grid update, so both are timed.  Lanes
  // keep their particles, or are re-assigned by rank within their cell
  // (layers).  The pair reduction is not a candidate: it can win on the
  // tuning state and then loses over the run on every workload measured
  // (1M snow 91.6 against 77.3 us, 16M 1.16 against 0.89 ms, the 256-scene
  // batch 1.43 against 1.13 ms per P2G); aemh_set_launch_shape selects it.
  // Where a node's values are many (the deterministic mode's 8 sums) or
  // cheap (velocity-only scenes, most of the set), lanes hitting the same node
  // cost more than the layering pass once particles cluster, which the
  // freshly sorted tuning state does not show: own lanes are charged 5 %
  // there (1M snow, deterministic mode: 1.49e9 when own lanes won the
  // tuning, 1.63e9 with layers; tier tick 0.271 against 0.258 ms).
  // The state being timed is often freshly sorted,
  // where a workgroup's particles span few blocks: a smaller tile budget
  // that wins there sends particles to the slow global path once the scene
  // spreads (1M snow: 3 tiles won at the tuning step, then took 106 us per
  // P2G against 82 us with 4).  Five FP32 tiles (4 sums, 6.9 KB each) still
  // leave LDS for 4 workgroups per CU, the VGPR limit, so that mode keeps the
  // full budget (1M snow P2G 81.5 -> 79 us against 4 tiles; 6 tiles cost a
  // workgroup per CU and 3 %).  The deterministic mode's tiles (8 sums,
  // 13.8 KB) do limit occupancy: it tries 2 to 4, each tile below 4 charged
  // 3 %.
  s = ensure_partials(c);
  if (s != AEMH_OK) return s;
  uint64_t vel_only_n = 0;
  for (const SceneDev& S : c->scd_h)
    if (S.vel_only) vel_only_n += S.count;
  const bool favor_layers = q32 || 2 * vel_only_n > (uint64_t)n;
  for (int layers = 0; layers <= 1; layers++)
  for (int part = 0; part <= 1; part++)
    for (int tiles = q32 ? 2 : kMaxTileBudget; tiles <= (q32 ? 4 : kMaxTileBudget); tiles++) {
      const int pairs = 0;
      ls.p2g_tiles = tiles;
      ls.p2g_partials = part;
      ls.p2g_pairs = pairs;
      ls.p2g_layers = layers;
      s = time_launches(c, kReps, &ms, [&] {
        launch_clear(st, c->sort, c->grid, acc_layout(c), ls.upd_wgs);
        if (q32) {
          det::launch_p2g(st, ls, a, c->scd_d, ns, c->mats_d, (int)c->nmat, n, c->sort, c->grid, storage);
          det::launch_grid_update(st, ls, c->scd_d, ns, c->sort, c->grid, n, c->nb_total);
        } else {
          fast::launch_p2g(st, ls, a, c->scd_d, ns, c->mats_d, (int)c->nmat, n, c->sort, c->grid, storage);
          fast::launch_grid_update(st, ls, c->scd_d, ns, c->sort, c->grid, n, c->nb_total);
        }
      });
      if (s != AEMH_OK) return s;
      const float score = ms * (1.0f + 0.03f * (float)((q32 ? 4 : kMaxTileBudget) - tiles)) *
                          (favor_layers && !layers ? 1.05f : 1.0f);
      if (score < best)
        best = score, best_ms = ms, pick = tiles, pick_part = part, pick_pairs = pairs, pick_layers = layers;
    }
  ls.p2g_tiles = pick;
  ls.p2g_partials = pick_part;
  ls.p2g_pairs = pick_pairs;
  ls.p2g_layers = pick_layers;
  c->tune_ms[0] = best_ms;
  // Partials in the chosen tile layout for the grid-update candidates.
  launch_clear(st, c->sort, c->grid, acc_layout(c), ls.upd_wgs);
  if (q32)
    det::launch_p2g(st, ls, a, c->scd_d, ns, c->mats_d, (int)c->nmat, n, c->sort, c->grid, storage);
  else
    fast::launch_p2g(st, ls, a, c->scd_d, ns, c->mats_d, (int)c->nmat, n, c->sort, c->grid, storage);
  best = 1e30f;
  pick = ls.upd_wgs;
  for (int wgs : {1024, 2048, 4096, 8192, 16384}) {
    ls.upd_wgs = wgs;
    s = time_launches(c, kReps, &ms, [&] {
      if (q32)
        det::launch_grid_update(st, ls, c->scd_d, ns, c->sort, c->grid, n, c->nb_total);
      else
        fast::launch_grid_update(st, ls, c->scd_d, ns, c->sort, c->grid, n, c->nb_total);
    });
    if (s != AEMH_OK) return s;
    if (ms < best) best = ms, pick = wgs;
  }
  ls.upd_wgs = pick;
  c->tune_ms[1] = best;
  best = 1e30f;
  pick = ls.g2p_wg;
  // Sizes other than 256 are charged 3 %: they tie or win narrowly on the
  // freshly sorted tuning state and lose as the scene spreads (1M snow: 128
  // picked in one run of four, 71.7 against 70.1 us per G2P over the timed
  // steps; 16M: 128 and 512 took 0.850 and 0.898 ms against 0.791).
  float best_score = 1e30f;
  for (int wg : {64, 128, 256, 512}) {
    ls.g2p_wg = wg;
    s = time_launches(c, kReps, &ms, [&] {
      if (q32)
        det::launch_g2p(st, ls, a, b, c->scd_d, ns, c->mats_d, (int)c->nmat, n, c->sort, c->grid, storage, 1, 0);
      else
        fast::launch_g2p(st, ls, a, b, c->scd_d, ns, c->mats_d, (int)c->nmat, n, c->sort, c->grid, storage, 1, 0);
    });
    if (s != AEMH_OK) return s;
    const float score = ms * (wg == 256 ? 1.0f : 1.03f);
    if (score < best_score) best_score = score, best = ms, pick = wg;
  }
  ls.g2p_wg = pick;
  c->tune_ms[2] = best;
  // The timed grid updates may have raised the Q32.32 range flag on sums
  // that no step will use; the next step recomputes it.
  HIPTRY(c, hipMemsetAsync(c->sort.counters + kCntOverflow, 0, 4, st));
  HIPTRY(c, hipStreamSynchronize(st));
  c->prof = prof;
  invalidate_graphs(c);
  c->tuned = true;
  return AEMH_OK;
}

aemh_status aemh_get_launch_shape(aemh_ctx* c, uint32_t* p2g_tiles, uint32_t* p2g_partials,
                                  uint32_t* g2p_wg, uint32_t* upd_wgs, float* tuned_ms3) {
  if (!c) return AEMH_ERR_ARG;
  if (p2g_tiles) *p2g_tiles = (uint32_t)c->shape.p2g_tiles;
  if (p2g_partials)
    *p2g_partials = (uint32_t)(c->shape.p2g_partials | (c->shape.p2g_pairs << 1) | (c->shape.p2g_layers << 2));
  if (g2p_wg) *g2p_wg = (uint32_t)c->shape.g2p_wg;
  if (upd_wgs) *upd_wgs = (uint32_t)c->shape.upd_wgs;
  if (tuned_ms3)
    for (int i = 0; i < 3; i++) tuned_ms3[i] = c->tuned ? c->tune_ms[i] : 0.0f;
  return AEMH_OK;
}

aemh_status aemh_set_launch_shape(aemh_ctx* c, uint32_t p2g_tiles, uint32_t p2g_partials,
                                  uint32_t g2p_wg, uint32_t upd_wgs) {
  if (!c) return AEMH_ERR_ARG;
  if (p2g_tiles < 2 || p2g_tiles > (uint32_t)kMaxTileBudget)
    return fail(c, AEMH_ERR_ARG, "p2g_tiles must be 2 to %d", kMaxTileBudget);
  if (p2g_partials > 7)
    return fail(c, AEMH_ERR_ARG, "p2g_partials must be 0-7 (bit 0 partial flush, bit 1 pairs, bit 2 layers)");
  if (g2p_wg != 64 && g2p_wg != 128 && g2p_wg != 256 && g2p_wg != 512)
    return fail(c, AEMH_ERR_ARG, "g2p_wg must be 64, 128, 256 or 512");
  if (upd_wgs < 1 || upd_wgs > 65536) return fail(c, AEMH_ERR_ARG, "upd_wgs must be in [1, 65536]");
  c->shape.p2g_tiles = (int)p2g_tiles;
  c->shape.p2g_partials = (int)(p2g_partials & 1u);
  c->shape.p2g_pairs = (int)((p2g_partials >> 1) & 1u);
  c->shape.p2g_layers = (int)((p2g_partials >> 2) & 1u);
  c->shape.g2p_wg = (int)g2p_wg;
  c->shape.upd_wgs = (int)upd_wgs;
  c->tuned = true;
  invalidate_graphs(c);
  return AEMH_OK;
}

aemh_status aemh_download_grid(aemh_ctx* c, uint32_t scene, float* mass, float* momentum3,
                               float* force3, float* vel3) {
  if (!c) return AEMH_ERR_ARG;
  if (scene >= c->scd_h.size()) return fail(c, AEMH_ERR_ARG, "no such scene");
  hipSetDevice(c->device);
  const SceneDev& S = c->scd_h[scene];
  const size_t nodes = (size_t)S.n[0] * S.n[1] * S.n[2];
  float* tmp = nullptr;
  HIPTRY(c, hipMalloc(&tmp, nodes * 10 * 4));
  HIPTRY(c, hipMemsetAsync(tmp, 0, nodes * 10 * 4, c->stream));
  launch_grid_readback(c->stream, c->scd_d, (int)scene, S, c->sort, c->grid, acc_layout(c), tmp,
                       tmp + nodes, tmp + 4 * nodes, tmp + 7 * nodes, nullptr, c->shape.upd_wgs);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess && mass) e = hipMemcpy(mass, tmp, nodes * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess && momentum3) e = hipMemcpy(momentum3, tmp + nodes, nodes * 12, hipMemcpyDeviceToHost);
  if (e == hipSuccess && force3) e = hipMemcpy(force3, tmp + 4 * nodes, nodes * 12, hipMemcpyDeviceToHost);
  if (e == hipSuccess && vel3) e = hipMemcpy(vel3, tmp + 7 * nodes, nodes * 12, hipMemcpyDeviceToHost);
  hipFree(tmp);
  if (e != hipSuccess) return fail(c, AEMH_ERR_HIP, "grid readback: %s", hipGetErrorName(e));
  return AEMH_OK;
}

aemh_status aemh_download_grid_q32(aemh_ctx* c, uint32_t scene, int64_t* q7) {
  if (!c || !q7) return fail(c, AEMH_ERR_ARG, "null argument");
  if (scene >= c->scd_h.size()) return fail(c, AEMH_ERR_ARG, "no such scene");
  if (c->cfg.accum != AEMH_ACCUM_Q32_32) return fail(c, AEMH_ERR_STATE, "not in Q32.32 mode");
  hipSetDevice(c->device);
  const SceneDev& S = c->scd_h[scene];
  const size_t nodes = (size_t)S.n[0] * S.n[1] * S.n[2];
  long long* tmp = nullptr;
  HIPTRY(c, hipMalloc(&tmp, nodes * 7 * 8));
  HIPTRY(c, hipMemsetAsync(tmp, 0, nodes * 7 * 8, c->stream));
  launch_grid_readback(c->stream, c->scd_d, (int)scene, S, c->sort, c->grid, kAccQ32, nullptr, nullptr,
                       nullptr, nullptr, tmp, c->shape.upd_wgs);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(q7, tmp, nodes * 7 * 8, hipMemcpyDeviceToHost);
  hipFree(tmp);
  if (e != hipSuccess) return fail(c, AEMH_ERR_HIP, "grid readback: %s", hipGetErrorName(e));
  return AEMH_OK;
}

aemh_status aemh_conservation_f64(aemh_ctx* c, double* out5) {
  if (!c || !out5) return fail(c, AEMH_ERR_ARG, "null argument");
  hipSetDevice(c->device);
  aemh_status s = ensure_work(c);
  if (s != AEMH_OK) return s;
  double* tmp = nullptr;
  HIPTRY(c, hipMalloc(&tmp, kConsScratch * sizeof(double)));
  launch_conservation(c->stream, c->work[c->cur].pb, c->n_work, tmp);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out5, tmp, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  hipFree(tmp);
  if (e != hipSuccess) return fail(c, AEMH_ERR_HIP, "reduction: %s", hipGetErrorName(e));
  return AEMH_OK;
}

// The same sums with no device allocation: the partials and the result go
// to this device's `scratch`, which holds what any working set needs
// (kConsScratch doubles, independent of the particle count), so no input
// exceeds it.  Every value there is written before it is read, so nothing
// is cleared first.  All calls on one device share that memory: calls that
// overlap on one device (two host threads, or two contexts of the device)
// race.  A working set that is not ready would have to be gathered, which
// allocates, so that case is refused.
aemh_status aemh_conservation_f64_noalloc(aemh_ctx* c, double* out5) {
  if (!c || !out5) return fail(c, AEMH_ERR_ARG, "null argument");
  static_assert(kConsScratch == 5 + 5 * kConsWgs, "scratch holds the sums and every partial");
  hipSetDevice(c->device);
  if (!c->cons_scratch)
    return fail(c, AEMH_ERR_HIP, "no device address for the reduction scratch: %s",
                c->scratch_err.c_str());
  if (!(c->work_valid && (c->idx_explicit || c->work_default || !c->mpss)))
    return fail(c, AEMH_ERR_STATE,
                "no current working set: step or set the indices first (gathering one allocates)");
  launch_conservation(c->stream, c->work[c->cur].pb, c->n_work, c->cons_scratch);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess)
    e = hipMemcpyAsync(out5, c->cons_scratch, 5 * sizeof(double), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(c, AEMH_ERR_HIP, "reduction: %s", hipGetErrorName(e));
  return AEMH_OK;
}

aemh_status aemh_svd_fp64_count(aemh_ctx* c, uint64_t* count) {
  if (!c || !count) return fail(c, AEMH_ERR_ARG, "null argument");
  hipSetDevice(c->device);
  uint32_t v = 0;
  if (c->sort.counters) {
    HIPTRY(c, hipStreamSynchronize(c->stream));
    HIPTRY(c, hipMemcpy(&v, c->sort.counters + kCntSvd64, 4, hipMemcpyDeviceToHost));
  }
  *count = v;
  return AEMH_OK;
}

aemh_status aemh_last_substep_info(aemh_ctx* c, uint32_t* active_blocks, uint32_t* particles) {
  if (!c) return AEMH_ERR_ARG;
  hipSetDevice(c->device);
  uint32_t nact = 0;
  if (c->sort.counters)
    HIPTRY(c, hipMemcpy(&nact, c->sort.counters + kCntActive, 4, hipMemcpyDeviceToHost));
  if (active_blocks) *active_blocks = nact;
  if (particles) *particles = c->n_work;
  return AEMH_OK;
}

aemh_status aemh_profile_enable(aemh_ctx* c, int on) {
  if (!c) return AEMH_ERR_ARG;
  c->prof = on != 0;
  return AEMH_OK;
}

aemh_status aemh_profile_get(aemh_ctx* c, const char* kernel, double* avg_ms, uint64_t* count) {
  if (!c || !kernel) return AEMH_ERR_ARG;
  auto it = c->prof_rec.find(kernel);
  if (it == c->prof_rec.end()) {
    if (avg_ms) *avg_ms = 0;
    if (count) *count = 0;
    return AEMH_OK;
  }
  if (avg_ms) *avg_ms = it->second.count ? it->second.total_ms / (double)it->second.count : 0.0;
  if (count) *count = it->second.count;
  return AEMH_OK;
}

aemh_status aemh_profile_reset(aemh_ctx* c) {
  if (!c) return AEMH_ERR_ARG;
  c->prof_rec.clear();
  return AEMH_OK;
}

// ---------------------------------------------------------------------------
// Resident MpssBuffer lifecycle (mpss.rs:161-297, 431-445)
// ---------------------------------------------------------------------------
aemh_status aemh_mpss_reserve(aemh_ctx* c, uint32_t capacity) {
  if (!c) return AEMH_ERR_ARG;
  if (capacity == 0) return fail(c, AEMH_ERR_ARG, "capacity must be > 0");
  hipSetDevice(c->device);
  aemh_status s = alloc_store(c, c->canon, capacity, 0);
  if (s != AEMH_OK) return s;
  c->canon.cap = capacity;
  if ((s = alloc_life(c, c->life, capacity)) != AEMH_OK) return s;
  // compaction targets are sized at the first compact() of this capacity
  for (void* p : {(void*)c->life_tmp.lifetime, (void*)c->life_tmp.age,
                  (void*)c->life_tmp.active, c->canon_tmp.slab})
    if (p) hipFree(p);
  c->life_tmp = MpssDev{};
  c->canon_tmp = ParticleStore{};
  PBuf& P = c->canon.pb;
  // MpssBuffer::new (mpss.rs:161-187): strain and C zero, jacobian and mass
  // 1, lifetime f32::MAX, age 0, material 0, nothing active.
  for (int a = 0; a < 3; a++) {
    launch_fill(c->stream, P.x[a], 0.0f, capacity);
    launch_fill(c->stream, P.v[a], 0.0f, capacity);
  }
  for (int k = 0; k < 9; k++) {
    launch_fill(c->stream, P.C[k], 0.0f, capacity);
    launch_fill(c->stream, P.F[k], 0.0f, capacity);
  }
  launch_fill(c->stream, P.J, 1.0f, capacity);
  launch_fill(c->stream, P.Jp, 1.0f, capacity);
  launch_fill(c->stream, P.mass, 1.0f, capacity);
  launch_fill(c->stream, c->life.lifetime, FLT_MAX, capacity);
  launch_fill(c->stream, c->life.age, 0.0f, capacity);
  HIPTRY(c, hipMemsetAsync(P.mat, 0, (size_t)capacity * 2, c->stream));
  HIPTRY(c, hipMemsetAsync(c->life.active, 0, capacity, c->stream));
  CHECK_LAUNCH(c);
  if (c->expired_d) hipFree(c->expired_d);
  c->expired_d = nullptr;
  HIPTRY(c, hipMalloc(&c->expired_d, ((size_t)capacity + 1) * 4));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  c->act_h.assign(capacity, 0);
  c->free_list.clear();
  c->mpss_count = 0;
  c->mpss = true;
  c->work_tier_key = 0;
  for (int t = 0; t < 3; t++) c->tier_n[t] = 0;
  c->work_valid = false;
  c->work_dirty = false;
  c->idx_explicit = false;
  c->idx_h.clear();
  if (c->sort.counters) HIPTRY(c, hipMemset(c->sort.counters, 0, 64));
  return AEMH_OK;
}

aemh_status aemh_mpss_spawn(aemh_ctx* c, uint32_t n, uint32_t* slots_out, uint32_t* n_spawned) {
  aemh_status s = mpss_ready(c);
  if (s != AEMH_OK) return s;
  if (n && !slots_out) return fail(c, AEMH_ERR_ARG, "null slot output");
  // spawn() n times (mpss.rs:189-203): the free list is popped last-in
  // first-out, then fresh slots at `count`; a full buffer yields None.
  std::vector<uint32_t> got;
  got.reserve(n);
  for (uint32_t i = 0; i < n; i++) {
    uint32_t idx = AEMH_NO_SLOT;
    if (!c->free_list.empty()) {
      idx = c->free_list.back();
      c->free_list.pop_back();
    } else if (c->mpss_count < c->canon.cap) {
      idx = c->mpss_count;
    }
    slots_out[i] = idx;
    if (idx == AEMH_NO_SLOT) continue;
    c->act_h[idx] = 1;
    c->mpss_count += 1;
    got.push_back(idx);
  }
  if (n_spawned) *n_spawned = (uint32_t)got.size();
  if (got.empty()) return AEMH_OK;
  if ((s = mpss_begin_write(c)) != AEMH_OK) return s;
  if ((s = upload_slots(c, got.data(), got.size())) != AEMH_OK) return s;
  launch_mpss_reset(c->stream, c->canon.pb, c->life, c->slots_d, (uint32_t)got.size());
  CHECK_LAUNCH(c);
  HIPTRY(c, hipStreamSynchronize(c->stream));
  return AEMH_OK;
}

aemh_status aemh_mpss_kill(aemh_ctx* c, const uint32_t* slots, uint32_t n) {
  aemh_status s = mpss_ready(c);
  if (s != AEMH_OK) return s;
  if (n && !slots) return fail(c, AEMH_ERR_ARG, "null slot list");
  return mpss_kill(c, slots, n);
}

aemh_status aemh_mpss_compact(aemh_ctx* c) {
  aemh_status s = mpss_ready(c);
  if (s != AEMH_OK) return s;
  const uint32_t cap = c->canon.cap;
  // compact() (mpss.rs:283-297): the reference's swap sequence, replayed on
  // slot labels, gives the permutation the device gather applies at once.
  std::vector<uint32_t> perm(cap);
  for (uint32_t i = 0; i < cap; i++) perm[i] = i;
  uint32_t write = 0;
  for (uint32_t read = 0; read < cap; read++) {
    if (!c->act_h[perm[read]]) continue;
    if (write != read) std::swap(perm[write], perm[read]);
    write++;
  }
  if ((s = mpss_begin_write(c)) != AEMH_OK) return s;
  if ((s = alloc_store(c, c->canon_tmp, cap, 0)) != AEMH_OK) return s;
  if (!c->life_tmp.lifetime && (s = alloc_life(c, c->life_tmp, cap)) != AEMH_OK) return s;
  if ((s = upload_slots(c, perm.data(), cap)) != AEMH_OK) return s;
  launch_mpss_gather(c->stream, c->canon.pb, c->life, c->canon_tmp.pb, c->life_tmp, c->slots_d, cap,
                     write);
  CHECK_LAUNCH(c);
  HIPTRY(c, hipStreamSynchronize(c->stream));
  std::swap(c->canon, c->canon_tmp);
  std::swap(c->life, c->life_tmp);
  c->canon.cap = cap;
  c->mpss_count = write;
  c->free_list.clear();
  for (uint32_t i = 0; i < cap; i++) c->act_h[i] = i < write ? 1 : 0;
  return AEMH_OK;
}

aemh_status aemh_mpss_update_lifetimes(aemh_ctx* c, float dt, uint32_t* expired_out,
                                       uint32_t max_out, uint32_t* n_expired) {
  aemh_status s = mpss_ready(c);
  if (s != AEMH_OK) return s;
  if (n_expired) *n_expired = 0;
  if (c->mpss_count == 0) return AEMH_OK;
  // update_lifetimes (mpss.rs:431-445): active slots below count age by dt;
  // those at or past their lifetime are killed in ascending slot order.
  HIPTRY(c, hipMemsetAsync(c->expired_d, 0, 4, c->stream));
  timed(c, "lifetimes", [&] {
    launch_mpss_age(c->stream, c->life, c->mpss_count, dt, c->expired_d + 1, c->expired_d);
  });
  CHECK_LAUNCH(c);
  uint32_t ne = 0;
  HIPTRY(c, hipMemcpyAsync(&ne, c->expired_d, 4, hipMemcpyDeviceToHost, c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  if (n_expired) *n_expired = ne;
  if (ne == 0) return AEMH_OK;
  std::vector<uint32_t> ex(ne);
  HIPTRY(c, hipMemcpy(ex.data(), c->expired_d + 1, (size_t)ne * 4, hipMemcpyDeviceToHost));
  std::sort(ex.begin(), ex.end());
  if (expired_out) memcpy(expired_out, ex.data(), (size_t)std::min(ne, max_out) * 4);
  return mpss_kill(c, ex.data(), ne);
}

aemh_status aemh_mpss_write(aemh_ctx* c, const uint32_t* slots, uint32_t n, const float* pos3,
                            const float* vel3, const float* C9, const float* F9, const float* J,
                            const float* Jp, const float* mass, const uint16_t* mat,
                            const float* lifetime, const float* age) {
  aemh_status s = mpss_ready(c);
  if (s != AEMH_OK) return s;
  if (n == 0) return AEMH_OK;
  if (!slots) return fail(c, AEMH_ERR_ARG, "null slot list");
  for (uint32_t i = 0; i < n; i++)
    if (slots[i] >= c->canon.cap)
      return fail(c, AEMH_ERR_ARG, "slot %u out of range (capacity %u)", slots[i], c->canon.cap);
  if ((s = mpss_begin_write(c)) != AEMH_OK) return s;
  if ((s = upload_slots(c, slots, n)) != AEMH_OK) return s;
  PBuf& P = c->canon.pb;
  struct Field {
    const float* h;
    int comps;
    float* const* d;
  };
  float* one[3] = {P.J, P.Jp, P.mass};
  const Field fields[] = {{pos3, 3, P.x},      {vel3, 3, P.v},     {C9, 9, P.C},
                          {F9, 9, P.F},        {J, 1, &one[0]},    {Jp, 1, &one[1]},
                          {mass, 1, &one[2]},  {lifetime, 1, &c->life.lifetime},
                          {age, 1, &c->life.age}};
  for (const Field& f : fields) {
    if (!f.h) continue;
    if ((s = ensure_staging(c, (size_t)n * f.comps * 4)) != AEMH_OK) return s;
    HIPTRY(c, hipMemcpyAsync(c->staging, f.h, (size_t)n * f.comps * 4, hipMemcpyHostToDevice,
                             c->stream));
    Ptr9 d{};
    for (int k = 0; k < f.comps; k++) d.p[k] = f.d[k];
    launch_rows_to_slots(c->stream, c->staging, d, f.comps, c->slots_d, n);
    CHECK_LAUNCH(c);
  }
  if (mat) {
    if ((s = ensure_staging(c, (size_t)n * 2)) != AEMH_OK) return s;
    HIPTRY(c, hipMemcpyAsync(c->staging, mat, (size_t)n * 2, hipMemcpyHostToDevice, c->stream));
    launch_u16_to_slots(c->stream, (const uint16_t*)c->staging, P.mat, c->slots_d, n);
    CHECK_LAUNCH(c);
  }
  HIPTRY(c, hipStreamSynchronize(c->stream));
  return AEMH_OK;
}

aemh_status aemh_mpss_read(aemh_ctx* c, const uint32_t* slots, uint32_t n, float* pos3, float* vel3,
                           float* C9, float* F9, float* J, float* Jp, float* mass, uint16_t* mat,
                           float* lifetime, float* age) {
  aemh_status s = mpss_ready(c);
  if (s != AEMH_OK) return s;
  if (n == 0) return AEMH_OK;
  if (!slots) return fail(c, AEMH_ERR_ARG, "null slot list");
  for (uint32_t i = 0; i < n; i++)
    if (slots[i] >= c->canon.cap)
      return fail(c, AEMH_ERR_ARG, "slot %u out of range (capacity %u)", slots[i], c->canon.cap);
  if ((s = flush_work(c)) != AEMH_OK) return s;
  if ((s = upload_slots(c, slots, n)) != AEMH_OK) return s;
  PBuf& P = c->canon.pb;
  struct Field {
    float* h;
    int comps;
    float* const* d;
  };
  float* one[3] = {P.J, P.Jp, P.mass};
  const Field fields[] = {{pos3, 3, P.x},      {vel3, 3, P.v},     {C9, 9, P.C},
                          {F9, 9, P.F},        {J, 1, &one[0]},    {Jp, 1, &one[1]},
                          {mass, 1, &one[2]},  {lifetime, 1, &c->life.lifetime},
                          {age, 1, &c->life.age}};
  for (const Field& f : fields) {
    if (!f.h) continue;
    if ((s = ensure_staging(c, (size_t)n * f.comps * 4)) != AEMH_OK) return s;
    Ptr9 d{};
    for (int k = 0; k < f.comps; k++) d.p[k] = f.d[k];
    launch_slots_to_rows(c->stream, d, c->staging, f.comps, c->slots_d, n);
    CHECK_LAUNCH(c);
    HIPTRY(c, hipMemcpyAsync(f.h, c->staging, (size_t)n * f.comps * 4, hipMemcpyDeviceToHost,
                             c->stream));
    HIPTRY(c, hipStreamSynchronize(c->stream));
  }
  if (mat) {
    if ((s = ensure_staging(c, (size_t)n * 2)) != AEMH_OK) return s;
    launch_slots_to_u16(c->stream, P.mat, (uint16_t*)c->staging, c->slots_d, n);
    CHECK_LAUNCH(c);
    HIPTRY(c, hipMemcpyAsync(mat, c->staging, (size_t)n * 2, hipMemcpyDeviceToHost, c->stream));
    HIPTRY(c, hipStreamSynchronize(c->stream));
  }
  return AEMH_OK;
}

aemh_status aemh_mpss_state(aemh_ctx* c, uint32_t* count, uint32_t* capacity, uint8_t* active) {
  aemh_status s = mpss_ready(c);
  if (s != AEMH_OK) return s;
  if (count) *count = c->mpss_count;
  if (capacity) *capacity = c->canon.cap;
  if (active) memcpy(active, c->act_h.data(), c->canon.cap);
  return AEMH_OK;
}

// ---------------------------------------------------------------------------
// LOD tier orchestrator (simulation.rs:480-700)
// ---------------------------------------------------------------------------
void aemh_lod_config_default(aemh_lod_config* L) {
  if (!L) return;
  memset(L, 0, sizeof(*L));
  // PhysicsLod::default (simulation.rs:35-48)
  L->near_field_distance = 50.0f;
  L->mid_field_distance = 200.0f;
  L->far_field_update_interval = 60;
  L->max_near_particles = 10000;
  L->mid_interval = 12;  // MID_INTERVAL (:586)
  // near grid: next_power_of_two(ceil(2 near / 1.6)) nodes of 1.6 (:219-227)
  L->near_dx = 1.6f;
  uint32_t n = (uint32_t)ceilf(2.0f * L->near_field_distance / L->near_dx);
  uint32_t p = 1;
  while (p < n) p <<= 1;
  L->near_n = p;
  // mid grid: 16 nodes spanning 2 mid (:234-241)
  L->mid_n = 16;
  L->mid_dx = (2.0f * L->mid_field_distance) / (float)L->mid_n;
}

aemh_status aemh_tier_set_config(aemh_ctx* c, const aemh_lod_config* L) {
  if (!c || !L) return fail(c, AEMH_ERR_ARG, "null argument");
  if (!(L->near_field_distance >= 0.0f) || !(L->mid_field_distance >= 0.0f) ||
      L->mid_interval == 0 || L->near_n == 0 || L->mid_n == 0 || !(L->near_dx > 0.0f) ||
      !(L->mid_dx > 0.0f))
    return fail(c, AEMH_ERR_ARG, "bad LOD configuration");
  c->lod = *L;
  c->lod_set = true;
  // A new manager: counters, cached player position and lists start over
  // (SimulationManager::new, simulation.rs:311-317).
  c->field_counter = 0;
  c->lod_counter = 0;
  for (int a = 0; a < 3; a++) c->cached_player[a] = 0.0f;
  for (int t = 0; t < 3; t++) c->tier_n[t] = 0;
  c->work_tier_key = 0;
  return AEMH_OK;
}

aemh_status aemh_tier_tick(aemh_ctx* c, const float player[3], float dt, aemh_tier_stats* out) {
  aemh_status s = mpss_ready(c);
  if (s != AEMH_OK) return s;
  if (!player) return fail(c, AEMH_ERR_ARG, "null player position");
  if (!c->lod_set) {
    aemh_lod_config_default(&c->lod);
    c->lod_set = true;
  }
  if (c->cfg.substeps == 0) return fail(c, AEMH_ERR_ARG, "substeps must be >= 1");
  aemh_tier_stats st{};
  c->field_counter += 1u;  // wrapping_add (:474)
  if (c->mpss_count == 0) {
    // Empty buffer (:692-697): gravity, integration and ageing visit active
    // slots only, and there are none.
    if (out) *out = st;
    return AEMH_OK;
  }
  if ((s = ensure_tier(c)) != AEMH_OK) return s;
  const aemh_lod_config& L = c->lod;
  // Reclassify when the player moved more than 5 m, the near cache is
  // empty, or every 8th tick (:488-494).
  const float mx = player[0] - c->cached_player[0];
  const float my = player[1] - c->cached_player[1];
  const float mz = player[2] - c->cached_player[2];
  const bool moved = mx * mx + my * my + mz * mz > 25.0f;
  c->lod_counter += 1u;
  if (moved || c->tier_n[0] == 0 || c->lod_counter % 8u == 0) {
    for (int a = 0; a < 3; a++) c->cached_player[a] = player[a];
    if ((s = tier_classify(c, player)) != AEMH_OK) return s;
    st.reclassified = 1;
  }
  // Moving-window origins: player - nx dx / 2 (:540-556).
  const float near_half = (float)L.near_n * L.near_dx * 0.5f;
  const float mid_half = (float)L.mid_n * L.mid_dx * 0.5f;
  for (int a = 0; a < 3; a++) {
    st.near_origin[a] = player[a] - near_half;
    st.mid_origin[a] = player[a] - mid_half;
  }
  const bool run_near = c->tier_n[0] > 0;
  const bool run_mid = c->field_counter % L.mid_interval == 0 && c->tier_n[1] > 0;
  const uint32_t far_interval = L.far_field_update_interval ? L.far_field_update_interval : 1u;
  const bool run_far = c->field_counter % far_interval == 0;
  if (run_near || run_mid) {
    // The near step and the mid tier's velocity-only step share no
    // particles and no grid, so both scenes advance in one batched launch
    // per kernel phase.
    const uint32_t* lists[2];
    uint32_t counts[2];
    int nl = 0;
    std::vector<aemh_scene> sc;
    const float sub = (float)c->cfg.substeps;
    if (run_near) {
      aemh_scene a{};
      a.first = 0;
      a.count = c->tier_n[0];
      a.nx = a.ny = a.nz = L.near_n;
      a.dx = L.near_dx;
      for (int k = 0; k < 3; k++) a.origin[k] = st.near_origin[k];
      a.dt = dt / sub;  // mpm_solver.rs:1075
      a.velocity_only = 0;
      a.clamp_mode = c->cfg.clamp_mode;
      sc.push_back(a);
      lists[nl] = c->tier_list[0];
      counts[nl++] = c->tier_n[0];
    }
    if (run_mid) {
      aemh_scene b{};
      b.first = run_near ? c->tier_n[0] : 0;
      b.count = c->tier_n[1];
      b.nx = b.ny = b.nz = L.mid_n;
      b.dx = L.mid_dx;
      for (int k = 0; k < 3; k++) b.origin[k] = st.mid_origin[k];
      const float mid_dt = dt * (float)L.mid_interval;  // :596
      b.dt = mid_dt / sub;                               // mpm_solver.rs:1347
      b.velocity_only = 1;
      b.clamp_mode = c->cfg.clamp_mode;
      sc.push_back(b);
      lists[nl] = c->tier_list[1];
      counts[nl++] = c->tier_n[1];
    }
    const uint64_t key = ((uint64_t)c->tier_version << 2) | (run_near ? 1u : 0u) |
                         (run_mid ? 2u : 0u);
    if ((s = tier_set_work(c, key, lists, counts, nl)) != AEMH_OK) return s;
    if ((s = set_scene_table(c, sc)) != AEMH_OK) return s;
    // Eager launches: the scene set changes from tick to tick.
    if ((s = run_substeps(c, c->cfg.substeps, false)) != AEMH_OK) return s;
  }
  if (run_far && c->tier_n[2]) {
    if ((s = flush_work(c)) != AEMH_OK) return s;
    timed(c, "far_update", [&] {
      launch_far_update(c->stream, c->canon.pb, c->tier_list[2], c->tier_n[2],
                        dt * (float)far_interval);
    });
    CHECK_LAUNCH(c);
    HIPTRY(c, hipStreamSynchronize(c->stream));
  }
  uint32_t ne = 0;
  // mpm_step_parallel_with_grid ages the buffer (mpm_solver.rs:1083), then
  // the tick does once more (:692).
  if (run_near) {
    if ((s = aemh_mpss_update_lifetimes(c, dt, nullptr, 0, &ne)) != AEMH_OK) return s;
    st.expired += ne;
  }
  if ((s = aemh_mpss_update_lifetimes(c, dt, nullptr, 0, &ne)) != AEMH_OK) return s;
  st.expired += ne;
  if (c->prof) collect_profile(c);
  st.near = c->tier_n[0];
  st.mid = c->tier_n[1];
  st.far = c->tier_n[2];
  st.near_ran = run_near;
  st.mid_ran = run_mid;
  st.far_ran = run_far;
  if (out) *out = st;
  return AEMH_OK;
}

aemh_status aemh_tier_lists(aemh_ctx* c, uint32_t* near, uint32_t* mid, uint32_t* far,
                            uint32_t* counts3) {
  aemh_status s = mpss_ready(c);
  if (s != AEMH_OK) return s;
  uint32_t* dst[3] = {near, mid, far};
  for (int t = 0; t < 3; t++) {
    if (counts3) counts3[t] = c->tier_n[t];
    if (dst[t] && c->tier_n[t])
      HIPTRY(c, hipMemcpy(dst[t], c->tier_list[t], (size_t)c->tier_n[t] * 4,
                          hipMemcpyDeviceToHost));
  }
  return AEMH_OK;
}

aemh_status aemh_tier_render(aemh_ctx* c, uint32_t tier, uint32_t take, float* pos3,
                             uint32_t* slots, uint32_t* n_out) {
  aemh_status s = mpss_ready(c);
  if (s != AEMH_OK) return s;
  if (tier > 2 || !n_out) return fail(c, AEMH_ERR_ARG, "tier must be 0..2 and n_out given");
  *n_out = 0;
  const uint32_t n = std::min(take, c->tier_n[tier]);
  if (n == 0) return AEMH_OK;
  if ((s = flush_work(c)) != AEMH_OK) return s;
  launch_live_filter(c->stream, c->tier_list[tier], n, c->life.active, c->mpss_count,
                     c->tier_cls);
  CHECK_LAUNCH(c);
  uint32_t n3[3] = {0, 0, 0};
  if ((s = tier_select(c, c->tier_cls, c->tier_list[tier], nullptr, n, c->tier_vals, nullptr,
                       nullptr, nullptr, n3)) != AEMH_OK)
    return s;
  const uint32_t m = n3[0];
  if (m && pos3) {
    if ((s = ensure_staging(c, (size_t)m * 12)) != AEMH_OK) return s;
    launch_gather_pos(c->stream, c->canon.pb, c->tier_vals, m, c->staging);
    CHECK_LAUNCH(c);
    HIPTRY(c, hipMemcpyAsync(pos3, c->staging, (size_t)m * 12, hipMemcpyDeviceToHost, c->stream));
  }
  if (m && slots)
    HIPTRY(c, hipMemcpyAsync(slots, c->tier_vals, (size_t)m * 4, hipMemcpyDeviceToHost, c->stream));
  HIPTRY(c, hipStreamSynchronize(c->stream));
  *n_out = m;
  return AEMH_OK;
}

}  // extern "C"
