// fixed_api.hip -- host side of the aemh_fixed_* entry points: the
// all-integer MpmSimulation (ae_physics/src/mpm.rs) resident on one gfx950
// device.  The host keeps the scalar state (time, step count, next id, the
// config fields the material models rewrite) and the per-substep constants;
// particles and grid live in HBM.  Calls are synchronous on return.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <string>
#include <vector>

#include "../../include/aemh.h"
#include "k_fixed.h"
#include "q32.h"

struct aemh_fixed {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  aemh_fixed_config cfg{};
  FxStore st;
  uint32_t n = 0, cap = 0;
  int64_t* grid = nullptr;
  uint32_t ncells = 0;
  int64_t* sin_tab = nullptr;
  int64_t* frag_off = nullptr;          // [9][9][3] fragment offsets for this cell size
  unsigned long long* bound = nullptr;  // current, tripped, max over substeps, spare
  unsigned long long* cell_bound = nullptr;  // [ncells], zero between substeps
  unsigned long long* counters = nullptr;    // substeps checked per cell, rebuilt serially
  uint64_t* key = nullptr;
  uint64_t* off = nullptr;
  int64_t* strain = nullptr;
  void* scan_tmp = nullptr;
  size_t scan_bytes = 0;
  uint32_t frac_cap = 0;
  uint64_t* total_h = nullptr;  // host-mapped
  uint64_t* total_d = nullptr;
  FxEvent* events = nullptr;
  uint32_t n_events = 0, ev_cap = 0;
  FxStatsAcc* acc = nullptr;
  uint64_t* stage = nullptr;
  size_t stage_words = 0;
  uint32_t* find_idx = nullptr;
  FxSortBufs sort{};
  uint32_t sort_cap = 0;
  uint32_t since_sort = 0;    // substeps since the P2G order was sorted
  uint32_t sort_every = 1;    // re-sort interval in substeps (aemh_fixed_set_sort_interval)
  bool order_valid = false;   // perm is a permutation of the current particles
  double time = 0.0;
  uint64_t step_count = 0, next_id = 1;
};

namespace {

using namespace q32;

constexpr uint32_t kMaxParticles = 1u << 30;
// Longest re-sort interval aemh_fixed_set_sort_interval accepts.
constexpr uint32_t kFxMaxSortEvery = 1u << 16;
// Below this many particles P2G scatters in particle order with one atomic
// per value: the cell sort's extra launches cost more than they save.
constexpr uint32_t kFxSortMin = 8192;

aemh_status fxfail(aemh_fixed* s, aemh_status code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (s) s->err = buf;
  return code;
}

#define FXTRY(s, expr)                                                                        \
  do {                                                                                        \
    hipError_t e_ = (expr);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fxfail(s, e_ == hipErrorOutOfMemory ? AEMH_ERR_OOM : AEMH_ERR_HIP, "%s: %s (%s)", \
                    #expr, hipGetErrorName(e_), hipGetErrorString(e_));                       \
  } while (0)

#define FXCHK(expr)                       \
  do {                                    \
    aemh_status st_ = (expr);             \
    if (st_ != AEMH_OK) return st_;       \
  } while (0)

// Rust `f64 as i64` (saturating, NaN -> 0) of value * 2^32: from_f32
int64_t from_f32(float v) {
  const double x = (double)v * 4294967296.0;
  if (x != x) return 0;
  if (x >= 9223372036854775808.0) return MAXV;
  if (x <= -9223372036854775808.0) return MINV;
  return (int64_t)x;
}
float to_f32(int64_t r) { return (float)r / 4294967296.0f; }
double to_f64(int64_t r) { return (double)r / 4294967296.0; }

const float kDefaultColor[4] = {0.8f, 0.6f, 0.4f, 1.0f};  // MpmParticle::default (mpm.rs:73-91)

uint64_t pack_color(float a, float b) {
  uint32_t x, y;
  memcpy(&x, &a, 4);
  memcpy(&y, &b, 4);
  return (uint64_t)x | ((uint64_t)y << 32);
}

// one particle record in plane order (k_fixed.h)
struct Rec {
  uint64_t w[FX_NPLANES];
};

Rec default_rec(uint64_t id) {
  Rec r;
  memset(&r, 0, sizeof(r));
  r.w[FX_F] = r.w[FX_F + 4] = r.w[FX_F + 8] = (uint64_t)ONE;
  r.w[FX_MASS] = r.w[FX_VOLUME] = (uint64_t)ONE;
  r.w[FX_TEMP] = (uint64_t)from_f32(293.0f);
  r.w[FX_FLAGS] = 1;
  r.w[FX_ID] = id;
  r.w[FX_COLOR] = pack_color(kDefaultColor[0], kDefaultColor[1]);
  r.w[FX_COLOR + 1] = pack_color(kDefaultColor[2], kDefaultColor[3]);
  return r;
}

size_t stride_for(uint32_t cap) { return (((size_t)cap + 63) & ~(size_t)63) + 288; }

aemh_status ensure_stage(aemh_fixed* s, size_t words) {
  if (words <= s->stage_words) return AEMH_OK;
  if (s->stage) FXTRY(s, hipFree(s->stage));
  s->stage = nullptr;
  s->stage_words = 0;
  FXTRY(s, hipMalloc((void**)&s->stage, words * 8));
  s->stage_words = words;
  return AEMH_OK;
}

aemh_status ensure_cap(aemh_fixed* s, uint64_t need) {
  if (need <= s->cap) return AEMH_OK;
  if (need > kMaxParticles) return fxfail(s, AEMH_ERR_ARG, "particle count %llu above %u", (unsigned long long)need, kMaxParticles);
  uint64_t nc = std::max<uint64_t>(need, std::max<uint64_t>(1024, (uint64_t)s->cap * 2));
  nc = std::min<uint64_t>(nc, kMaxParticles);
  FxStore ns;
  ns.stride = stride_for((uint32_t)nc);
  FXTRY(s, hipMalloc((void**)&ns.base, ns.stride * FX_NPLANES * 8));
  if (s->st.base) {
    if (s->n)
      FXTRY(s, hipMemcpy2DAsync(ns.base, ns.stride * 8, s->st.base, s->st.stride * 8, (size_t)s->n * 8, FX_NPLANES,
                                hipMemcpyDeviceToDevice, s->stream));
    FXTRY(s, hipStreamSynchronize(s->stream));
    FXTRY(s, hipFree(s->st.base));
  }
  s->st = ns;
  s->cap = (uint32_t)nc;
  return AEMH_OK;
}

// uploads m records to [at, at + m)
aemh_status put_records(aemh_fixed* s, uint32_t at, const Rec* recs, uint32_t m) {
  if (!m) return AEMH_OK;
  s->order_valid = false;
  FXCHK(ensure_stage(s, (size_t)m * FX_NPLANES));
  FXTRY(s, hipMemcpyAsync(s->stage, recs, (size_t)m * sizeof(Rec), hipMemcpyHostToDevice, s->stream));
  FXTRY(s, fx_launch_put(s->st, at, s->stage, m, s->stream));
  FXTRY(s, hipStreamSynchronize(s->stream));
  return AEMH_OK;
}

aemh_status get_records(aemh_fixed* s, uint32_t first, uint32_t m, Rec* recs) {
  if (!m) return AEMH_OK;
  FXCHK(ensure_stage(s, (size_t)m * FX_NPLANES));
  FXTRY(s, fx_launch_get(s->st, first, m, s->stage, s->stream));
  FXTRY(s, hipMemcpyAsync(recs, s->stage, (size_t)m * sizeof(Rec), hipMemcpyDeviceToHost, s->stream));
  FXTRY(s, hipStreamSynchronize(s->stream));
  return AEMH_OK;
}

// init_particles (mpm.rs:179-206)
aemh_status init_particles(aemh_fixed* s) {
  const uint64_t count = s->cfg.particle_count;
  if (count == 0) return AEMH_OK;
  const float ppd_f = ceilf(cbrtf((float)count));
  const uint32_t ppd = ppd_f >= 4294967296.0f ? UINT32_MAX : (ppd_f > 0.0f ? (uint32_t)ppd_f : 0u);
  const uint64_t total = (uint64_t)ppd * ppd * ppd;
  if (total > kMaxParticles) return fxfail(s, AEMH_ERR_ARG, "init_particles: %llu particles", (unsigned long long)total);
  const int64_t cell = s->cfg.cell_size;
  const int64_t spacing = q32::div(cell, from_int((int32_t)ppd));
  const int64_t lift = mul(mul(from_int((int32_t)s->cfg.grid_resolution[1]), cell), from_f32(0.3f));
  const int64_t vol = mul(mul(spacing, spacing), spacing);
  const int64_t mass = mul(s->cfg.density, vol);
  FXCHK(ensure_cap(s, total));
  std::vector<Rec> recs;
  recs.reserve((size_t)std::min<uint64_t>(total, 1u << 20));
  for (uint32_t i = 0; i < ppd; ++i)
    for (uint32_t j = 0; j < ppd; ++j)
      for (uint32_t k = 0; k < ppd; ++k) {
        Rec r = default_rec(s->next_id++);
        r.w[FX_POS] = (uint64_t)mul(add(from_int((int32_t)i), HALF), spacing);
        r.w[FX_POS + 1] = (uint64_t)add(mul(add(from_int((int32_t)j), HALF), spacing), lift);
        r.w[FX_POS + 2] = (uint64_t)mul(add(from_int((int32_t)k), HALF), spacing);
        r.w[FX_MASS] = (uint64_t)mass;
        r.w[FX_VOLUME] = (uint64_t)vol;
        recs.push_back(r);
        if (recs.size() == (1u << 20)) {
          FXCHK(put_records(s, s->n, recs.data(), (uint32_t)recs.size()));
          s->n += (uint32_t)recs.size();
          recs.clear();
        }
      }
  FXCHK(put_records(s, s->n, recs.data(), (uint32_t)recs.size()));
  s->n += (uint32_t)recs.size();
  return AEMH_OK;
}

// fragment offsets of handle_fracture (mpm.rs:679-689) for 1..9 fragments;
// the f32 sin/cos are the C library's, as Rust's f32::sin/cos on Linux
aemh_status upload_fragment_offsets(aemh_fixed* s) {
  const int64_t cell = s->cfg.cell_size;
  const int64_t half = mul(cell, from_f32(0.5f));
  const int64_t fifth = from_f32(0.2f);
  const float tau = 6.28318530717958647692f;
  std::vector<int64_t> t((size_t)kFxFragTable * kFxFragTable * 3, 0);
  for (int nf = 1; nf <= kFxFragTable; ++nf)
    for (int fi = 0; fi < nf; ++fi) {
      volatile float fi_f = (float)fi;  // keep the library calls at run time
      const float angle = fi_f * tau / (float)nf;
      int64_t* o = &t[((size_t)(nf - 1) * kFxFragTable + fi) * 3];
      o[0] = mul(from_f32(cosf(angle)), half);
      o[1] = mul(mul(from_f32(sinf(angle)), cell), fifth);
      o[2] = mul(from_f32(sinf(fi_f * 1.7f)), half);
    }
  FXTRY(s, hipMemcpy(s->frag_off, t.data(), t.size() * 8, hipMemcpyHostToDevice));
  return AEMH_OK;
}

// SIN_TABLE (fixed_point.rs:874-883)
aemh_status upload_sin_table(aemh_fixed* s) {
  std::vector<int64_t> t(1 << kSinBits);
  const double half_pi = 1.57079632679489661923;
  for (int i = 0; i < (1 << kSinBits); ++i) {
    volatile double x = (double)i / (double)(1 << kSinBits) * half_pi;
    const double v = ::sin(x) * 4294967296.0;
    t[i] = v >= 9223372036854775808.0 ? MAXV : (int64_t)v;
  }
  FXTRY(s, hipMemcpy(s->sin_tab, t.data(), t.size() * 8, hipMemcpyHostToDevice));
  return AEMH_OK;
}

aemh_status ensure_frac(aemh_fixed* s, uint32_t n) {
  if (n <= s->frac_cap) return AEMH_OK;
  const uint32_t nc = std::max<uint32_t>(n, std::max<uint32_t>(1024, s->frac_cap * 2));
  if (s->key) FXTRY(s, hipFree(s->key));
  if (s->off) FXTRY(s, hipFree(s->off));
  if (s->strain) FXTRY(s, hipFree(s->strain));
  if (s->scan_tmp) FXTRY(s, hipFree(s->scan_tmp));
  s->key = s->off = nullptr;
  s->strain = nullptr;
  s->scan_tmp = nullptr;
  s->frac_cap = 0;
  FXTRY(s, hipMalloc((void**)&s->key, (size_t)nc * 8));
  FXTRY(s, hipMalloc((void**)&s->off, (size_t)nc * 8));
  FXTRY(s, hipMalloc((void**)&s->strain, (size_t)nc * 8));
  s->scan_bytes = std::max<size_t>(fx_scan_temp_bytes(nc), 256);
  FXTRY(s, hipMalloc(&s->scan_tmp, s->scan_bytes));
  s->frac_cap = nc;
  return AEMH_OK;
}

aemh_status ensure_events(aemh_fixed* s, uint64_t need) {
  if (need <= s->ev_cap) return AEMH_OK;
  if (need > UINT32_MAX) return fxfail(s, AEMH_ERR_ARG, "too many fracture events");
  const uint32_t nc = (uint32_t)std::max<uint64_t>(need, std::max<uint64_t>(256, (uint64_t)s->ev_cap * 2));
  FxEvent* ne = nullptr;
  FXTRY(s, hipMalloc((void**)&ne, (size_t)nc * sizeof(FxEvent)));
  if (s->events) {
    if (s->n_events)
      FXTRY(s, hipMemcpyAsync(ne, s->events, (size_t)s->n_events * sizeof(FxEvent), hipMemcpyDeviceToDevice, s->stream));
    FXTRY(s, hipStreamSynchronize(s->stream));
    FXTRY(s, hipFree(s->events));
  }
  s->events = ne;
  s->ev_cap = nc;
  return AEMH_OK;
}

// handle_fracture (mpm.rs:641-716)
aemh_status fracture(aemh_fixed* s) {
  const uint32_t n0 = s->n;
  if (!n0) return AEMH_OK;
  FXCHK(ensure_frac(s, n0));
  FXTRY(s, fx_launch_frac_mark(s->st, n0, s->cfg.fracture_strain, s->key, s->strain, s->stream));
  FXTRY(s, fx_launch_scan(s->key, s->off, n0, s->scan_tmp, s->scan_bytes, s->total_d, s->stream));
  FXTRY(s, hipStreamSynchronize(s->stream));
  const uint64_t total = *s->total_h;
  const uint32_t nev = (uint32_t)(total >> 32), nfr = (uint32_t)total;
  if (!nev) return AEMH_OK;
  FXCHK(ensure_cap(s, (uint64_t)n0 + nfr));
  FXCHK(ensure_events(s, (uint64_t)s->n_events + nev));
  FXTRY(s, fx_launch_frac_spawn(s->st, n0, s->key, s->off, s->strain, s->frag_off, s->next_id, s->time,
                                s->events + s->n_events, s->stream));
  s->n = n0 + nfr;
  s->next_id += nfr;
  s->n_events += nev;
  s->order_valid = false;
  return AEMH_OK;
}

aemh_status ensure_sort(aemh_fixed* s, uint32_t n) {
  if (!s->sort.bins) {  // bins follow the grid, fixed at create
    s->sort.nbins = (uint32_t)((uint64_t)(s->cfg.grid_resolution[0] + 2) * (s->cfg.grid_resolution[1] + 2) *
                               (s->cfg.grid_resolution[2] + 2)) + 1;
    FXTRY(s, hipMalloc((void**)&s->sort.bins, (size_t)s->sort.nbins * 4));
    FXTRY(s, hipMalloc((void**)&s->sort.cursor, (size_t)s->sort.nbins * 4));
    s->sort.temp_bytes = std::max<size_t>(fx_sort_temp_bytes(s->sort.nbins), 256);
    FXTRY(s, hipMalloc(&s->sort.temp, s->sort.temp_bytes));
  }
  if (n <= s->sort_cap) return AEMH_OK;
  s->order_valid = false;
  const uint32_t nc = std::max<uint32_t>(n, std::max<uint32_t>(1024, s->sort_cap * 2));
  for (int k = 0; k < 2; ++k) {
    void* old[] = {s->sort.keys[k], s->sort.rec[k], s->sort.perm[k]};
    for (void* b : old)
      if (b) FXTRY(s, hipFree(b));
    s->sort.keys[k] = nullptr;
    s->sort.rec[k] = nullptr;
    s->sort.perm[k] = nullptr;
  }
  if (s->sort.rank) FXTRY(s, hipFree(s->sort.rank));
  s->sort.rank = nullptr;
  s->sort_cap = 0;
  for (int k = 0; k < 2; ++k) {
    FXTRY(s, hipMalloc((void**)&s->sort.keys[k], (size_t)nc * 4));
    FXTRY(s, hipMalloc((void**)&s->sort.rec[k], (size_t)nc * kFxRecWords * 8));
    FXTRY(s, hipMalloc((void**)&s->sort.perm[k], (size_t)nc * 4));
  }
  FXTRY(s, hipMalloc((void**)&s->sort.rank, (size_t)nc * 4));
  s->sort_cap = nc;
  return AEMH_OK;
}

// substep (mpm.rs:253-263)
aemh_status substep(aemh_fixed* s, int64_t sdt) {
  aemh_fixed_config& c = s->cfg;
  const uint32_t model = c.material_model;
  if (model == AEMH_FX_SNOW) c.hardening = c.model_param;  // :466-469, before the loop
  FxSubstep sp;
  memset(&sp, 0, sizeof(sp));
  sp.cell = c.cell_size;
  sp.inv_cell = q32::div(ONE, c.cell_size);
  sp.sdt = sdt;
  for (int a = 0; a < 3; ++a) {
    sp.gdt[a] = mul(c.gravity[a], sdt);
    sp.res[a] = (int32_t)c.grid_resolution[a];
  }
  sp.mu = q32::div(c.youngs_modulus, mul(TWO, add(ONE, c.poissons_ratio)));
  sp.lambda = q32::div(mul(c.youngs_modulus, c.poissons_ratio),
                  mul(add(ONE, c.poissons_ratio), sub(ONE, mul(TWO, c.poissons_ratio))));
  sp.E = c.youngs_modulus;
  sp.yield = c.yield_stress;
  sp.hardening = c.hardening;
  if (model == AEMH_FX_GRANULAR) {  // :505-506
    const float rads_per_deg = 3.14159265358979323846f / 180.0f;
    volatile float deg = to_f32(c.model_param);
    sp.mu_s = from_f32(tanf(deg * rads_per_deg));
  }
  sp.viscosity = c.model_param;
  sp.j_lo = from_f32(0.01f);
  sp.j_hi = from_f32(100.0f);
  sp.crit_c = sub(ONE, from_f32(2.5e-2f));
  sp.crit_s = add(ONE, from_f32(7.5e-3f));
  sp.n = s->n;
  FXTRY(s, hipMemsetAsync(s->grid, 0, (size_t)s->ncells * 32, s->stream));
  int resort = FX_ORDER_KEEP;
  const bool sorted = s->n >= kFxSortMin;
  if (sorted) {
    FXCHK(ensure_sort(s, s->n));
    resort = !s->order_valid ? FX_ORDER_FROM_PLANES
                             : (s->since_sort >= s->sort_every ? FX_ORDER_RESORT : FX_ORDER_KEEP);
    s->since_sort = resort != FX_ORDER_KEEP ? 1 : s->since_sort + 1;
    s->order_valid = true;
  } else {
    s->order_valid = false;  // G2P does not refresh the records on this path
  }
  FXTRY(s, fx_launch_p2g(s->st, sp, s->grid, s->bound, s->cell_bound, s->ncells, s->counters,
                         sorted ? &s->sort : nullptr, resort, s->stream));
  FXTRY(s, fx_launch_grid(s->grid, s->ncells, sp.gdt, s->bound, s->stream));
  FXTRY(s, fx_launch_g2p((int)model, s->st, sp, s->grid, s->sin_tab, sorted ? &s->sort : nullptr, s->stream));
  if (model == AEMH_FX_BRITTLE) {  // :462-465
    c.fracture_strain = c.model_param;
    c.enable_fracture = 1;
  }
  if (c.enable_fracture) FXCHK(fracture(s));
  return AEMH_OK;
}

aemh_status check_records(aemh_fixed* s, uint64_t first, uint64_t count) {
  if (first > s->n || count > s->n - first) return fxfail(s, AEMH_ERR_ARG, "records [%llu, +%llu) outside %u particles",
                                                          (unsigned long long)first, (unsigned long long)count, s->n);
  return AEMH_OK;
}

}  // namespace

extern "C" {

void aemh_fixed_config_default(aemh_fixed_config* c) {
  if (!c) return;
  memset(c, 0, sizeof(*c));
  for (int a = 0; a < 3; ++a) c->grid_resolution[a] = 64;
  c->cell_size = from_f32(0.1f);
  c->particle_count = 10000;
  c->youngs_modulus = from_f32(1e6f);
  c->poissons_ratio = from_f32(0.3f);
  c->yield_stress = from_f32(1e4f);
  c->hardening = from_f32(0.1f);
  c->density = from_f32(1000.0f);
  c->gravity[0] = from_f32(0.0f);
  c->gravity[1] = from_f32(-9.81f);
  c->gravity[2] = from_f32(0.0f);
  c->dt = from_f32(1.0f / 240.0f);
  c->substeps = 4;
  c->material_model = AEMH_FX_ELASTOPLASTIC;
  c->model_param = 0;
  c->enable_fracture = 0;
  c->fracture_strain = from_f32(0.15f);
}

aemh_status aemh_fixed_create(aemh_fixed** out, int device, const aemh_fixed_config* cfg) {
  if (!out || !cfg) return AEMH_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return AEMH_ERR_NO_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) return AEMH_ERR_NO_DEVICE;
  if (strncmp(prop.gcnArchName, "gfx950", 6) != 0) return AEMH_ERR_NO_DEVICE;
  const uint64_t cells = (uint64_t)cfg->grid_resolution[0] * cfg->grid_resolution[1] * cfg->grid_resolution[2];
  if (cells == 0 || cells > (1ull << 31) || cfg->grid_resolution[0] > INT32_MAX ||
      cfg->grid_resolution[1] > INT32_MAX || cfg->grid_resolution[2] > INT32_MAX || cfg->substeps > INT32_MAX ||
      cfg->material_model > AEMH_FX_SNOW ||
      (uint64_t)(cfg->grid_resolution[0] + 2ull) * (cfg->grid_resolution[1] + 2ull) * (cfg->grid_resolution[2] + 2ull) >=
          (1ull << 32) - 128)  // cell keys stay clear of the idle-lane keys near 2^32
    return AEMH_ERR_ARG;
  aemh_fixed* s = new aemh_fixed();
  s->device = device;
  s->cfg = *cfg;
  s->ncells = (uint32_t)cells;
  aemh_status st = AEMH_OK;
  auto setup = [&]() -> aemh_status {
    FXTRY(s, hipSetDevice(device));
    FXTRY(s, hipStreamCreate(&s->stream));
    FXTRY(s, hipMalloc((void**)&s->grid, (size_t)cells * 32));
    FXTRY(s, hipMalloc((void**)&s->sin_tab, (size_t)(1 << kSinBits) * 8));
    FXTRY(s, hipMalloc((void**)&s->frag_off, (size_t)kFxFragTable * kFxFragTable * 3 * 8));
    FXTRY(s, hipMalloc((void**)&s->bound, 4 * sizeof(unsigned long long)));
    FXTRY(s, hipMemset(s->bound, 0, 4 * sizeof(unsigned long long)));
    FXTRY(s, hipMalloc((void**)&s->cell_bound, (size_t)cells * sizeof(unsigned long long)));
    FXTRY(s, hipMemset(s->cell_bound, 0, (size_t)cells * sizeof(unsigned long long)));
    FXTRY(s, hipMalloc((void**)&s->counters, 2 * sizeof(unsigned long long)));
    FXTRY(s, hipMemset(s->counters, 0, 2 * sizeof(unsigned long long)));
    FXTRY(s, hipMalloc((void**)&s->acc, sizeof(FxStatsAcc)));
    FXTRY(s, hipMalloc((void**)&s->find_idx, 4));
    FXTRY(s, hipHostMalloc((void**)&s->total_h, 8, hipHostMallocMapped | hipHostMallocCoherent));
    *s->total_h = 0;
    FXTRY(s, hipHostGetDevicePointer((void**)&s->total_d, s->total_h, 0));
    FXCHK(upload_sin_table(s));
    FXCHK(upload_fragment_offsets(s));
    FXCHK(ensure_cap(s, std::max<uint64_t>(1024, cfg->particle_count)));
    return init_particles(s);
  };
  st = setup();
  if (st != AEMH_OK) {
    fprintf(stderr, "aemh_fixed_create: %s\n", s->err.c_str());
    aemh_fixed_destroy(s);
    return st;
  }
  *out = s;
  return AEMH_OK;
}

void aemh_fixed_destroy(aemh_fixed* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  void* bufs[] = {s->st.base, s->grid, s->sin_tab, s->frag_off, s->bound, s->cell_bound, s->counters,
                  s->key, s->off, s->strain, s->scan_tmp, s->events, s->acc, s->stage, s->find_idx,
                  s->sort.keys[0], s->sort.keys[1], s->sort.rec[0], s->sort.rec[1], s->sort.perm[0],
                  s->sort.perm[1], s->sort.rank, s->sort.bins, s->sort.cursor, s->sort.temp};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  if (s->total_h) (void)hipHostFree(s->total_h);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

const char* aemh_fixed_last_error(const aemh_fixed* s) { return s ? s->err.c_str() : "null handle"; }

aemh_status aemh_fixed_add_particles(aemh_fixed* s, uint64_t n, const int64_t* pos3, const int64_t* vel3,
                                     const int64_t* mass, const float* color4, uint64_t* first_id_out) {
  if (!s) return AEMH_ERR_ARG;
  if (n && (!pos3 || !vel3 || !mass)) return fxfail(s, AEMH_ERR_ARG, "add_particles: null array");
  (void)hipSetDevice(s->device);
  FXCHK(ensure_cap(s, (uint64_t)s->n + n));
  if (first_id_out) *first_id_out = s->next_id;
  const uint64_t chunk = 1u << 20;
  std::vector<Rec> recs;
  for (uint64_t b = 0; b < n; b += chunk) {
    const uint64_t m = std::min(chunk, n - b);
    recs.resize((size_t)m);
    for (uint64_t q = 0; q < m; ++q) {
      const uint64_t i = b + q;
      Rec& r = recs[(size_t)q];
      r = default_rec(s->next_id++);
      for (int a = 0; a < 3; ++a) {
        r.w[FX_POS + a] = (uint64_t)pos3[i * 3 + a];
        r.w[FX_VEL + a] = (uint64_t)vel3[i * 3 + a];
      }
      r.w[FX_MASS] = (uint64_t)mass[i];
      r.w[FX_VOLUME] = (uint64_t)q32::div(mass[i], s->cfg.density);  // mass / density (:224)
      if (color4) {
        r.w[FX_COLOR] = pack_color(color4[i * 4], color4[i * 4 + 1]);
        r.w[FX_COLOR + 1] = pack_color(color4[i * 4 + 2], color4[i * 4 + 3]);
      }
    }
    FXCHK(put_records(s, s->n, recs.data(), (uint32_t)m));
    s->n += (uint32_t)m;
  }
  return AEMH_OK;
}

aemh_status aemh_fixed_add_particle(aemh_fixed* s, const int64_t* pos3, const int64_t* vel3, int64_t mass,
                                    const float* color4, uint64_t* id_out) {
  return aemh_fixed_add_particles(s, 1, pos3, vel3, &mass, color4, id_out);
}

aemh_status aemh_fixed_remove_particle(aemh_fixed* s, uint64_t id, int* removed) {
  if (!s) return AEMH_ERR_ARG;
  (void)hipSetDevice(s->device);
  if (removed) *removed = 0;
  if (!s->n) return AEMH_OK;
  FXTRY(s, hipMemsetAsync(s->find_idx, 0xff, 4, s->stream));
  FXTRY(s, fx_launch_find(s->st, s->n, id, s->find_idx, s->stream));
  uint32_t idx = UINT32_MAX;
  FXTRY(s, hipMemcpyAsync(&idx, s->find_idx, 4, hipMemcpyDeviceToHost, s->stream));
  FXTRY(s, hipStreamSynchronize(s->stream));
  if (idx == UINT32_MAX) return AEMH_OK;
  // Vec::remove: the tail moves down one slot, order kept, in chunks from
  // the front (each chunk reads the slots just past those it writes)
  const uint32_t tail = s->n - idx - 1;
  const uint32_t chunk = 1u << 20;
  for (uint32_t done = 0; done < tail;) {
    const uint32_t m = std::min(chunk, tail - done);
    FXCHK(ensure_stage(s, (size_t)m * FX_NPLANES));
    FXTRY(s, fx_launch_get(s->st, idx + 1 + done, m, s->stage, s->stream));
    FXTRY(s, fx_launch_put(s->st, idx + done, s->stage, m, s->stream));
    done += m;
  }
  FXTRY(s, hipStreamSynchronize(s->stream));
  s->n -= 1;
  s->order_valid = false;
  if (removed) *removed = 1;
  return AEMH_OK;
}

aemh_status aemh_fixed_step(aemh_fixed* s, uint32_t n_steps) {
  if (!s) return AEMH_ERR_ARG;
  (void)hipSetDevice(s->device);
  for (uint32_t k = 0; k < n_steps; ++k) {
    // step (mpm.rs:244-251)
    const int64_t sdt = q32::div(s->cfg.dt, from_int((int32_t)s->cfg.substeps));
    for (uint32_t q = 0; q < s->cfg.substeps; ++q) FXCHK(substep(s, sdt));
    s->time += to_f64(s->cfg.dt);
    s->step_count += 1;
  }
  FXTRY(s, hipStreamSynchronize(s->stream));
  return AEMH_OK;
}

aemh_status aemh_fixed_set_material_model(aemh_fixed* s, uint32_t model, int64_t model_param) {
  if (!s) return AEMH_ERR_ARG;
  if (model > AEMH_FX_SNOW) return fxfail(s, AEMH_ERR_ARG, "material model %u", model);
  s->cfg.material_model = model;
  s->cfg.model_param = model_param;
  return AEMH_OK;
}

aemh_status aemh_fixed_set_sort_interval(aemh_fixed* s, uint32_t substeps) {
  if (!s) return AEMH_ERR_ARG;
  if (substeps < 1 || substeps > kFxMaxSortEvery)
    return fxfail(s, AEMH_ERR_ARG, "sort interval %u outside [1, %u]", substeps, kFxMaxSortEvery);
  s->sort_every = substeps;
  return AEMH_OK;
}

aemh_status aemh_fixed_enable_fracture(aemh_fixed* s, int64_t threshold) {
  if (!s) return AEMH_ERR_ARG;
  s->cfg.enable_fracture = 1;
  s->cfg.fracture_strain = threshold;
  return AEMH_OK;
}

aemh_status aemh_fixed_disable_fracture(aemh_fixed* s) {
  if (!s) return AEMH_ERR_ARG;
  s->cfg.enable_fracture = 0;
  return AEMH_OK;
}

aemh_status aemh_fixed_get_state(aemh_fixed* s, aemh_fixed_state* o) {
  if (!s || !o) return AEMH_ERR_ARG;
  memset(o, 0, sizeof(*o));
  (void)hipSetDevice(s->device);
  unsigned long long cnt[2], b[4];
  FXTRY(s, hipMemcpyAsync(cnt, s->counters, sizeof(cnt), hipMemcpyDeviceToHost, s->stream));
  FXTRY(s, hipMemcpyAsync(b, s->bound, sizeof(b), hipMemcpyDeviceToHost, s->stream));
  FXTRY(s, hipStreamSynchronize(s->stream));
  o->checked_substeps = cnt[0];
  o->serial_substeps = cnt[1];
  o->max_scatter_bound = b[3] ? UINT64_MAX : b[2];
  o->particles = s->n;
  o->step_count = s->step_count;
  o->time = s->time;
  o->next_particle_id = s->next_id;
  o->fracture_events = s->n_events;
  o->hardening = s->cfg.hardening;
  o->fracture_strain = s->cfg.fracture_strain;
  o->enable_fracture = s->cfg.enable_fracture;
  o->material_model = s->cfg.material_model;
  o->model_param = s->cfg.model_param;
  return AEMH_OK;
}

aemh_status aemh_fixed_read_particles(aemh_fixed* s, uint64_t first, uint64_t count, int64_t* q29, uint8_t* flags3,
                                      uint64_t* ids9, float* color4) {
  if (!s) return AEMH_ERR_ARG;
  FXCHK(check_records(s, first, count));
  (void)hipSetDevice(s->device);
  const uint64_t chunk = 1u << 20;
  std::vector<Rec> recs;
  for (uint64_t b = 0; b < count; b += chunk) {
    const uint64_t m = std::min(chunk, count - b);
    recs.resize((size_t)m);
    FXCHK(get_records(s, (uint32_t)(first + b), (uint32_t)m, recs.data()));
    for (uint64_t q = 0; q < m; ++q) {
      const Rec& r = recs[(size_t)q];
      const uint64_t i = b + q;
      if (q29)
        for (int k = 0; k < kFxScalarPlanes; ++k) q29[i * 29 + k] = (int64_t)r.w[k];
      if (flags3) {
        flags3[i * 3] = (uint8_t)(r.w[FX_FLAGS] & 1);
        flags3[i * 3 + 1] = (uint8_t)((r.w[FX_FLAGS] >> 8) & 1);
        flags3[i * 3 + 2] = (uint8_t)(r.w[FX_FLAGS] >> 16);
      }
      if (ids9)
        for (int k = 0; k < 9; ++k) ids9[i * 9 + k] = r.w[FX_ID + k];
      if (color4) {
        const uint32_t c[4] = {(uint32_t)r.w[FX_COLOR], (uint32_t)(r.w[FX_COLOR] >> 32), (uint32_t)r.w[FX_COLOR + 1],
                               (uint32_t)(r.w[FX_COLOR + 1] >> 32)};
        memcpy(color4 + i * 4, c, 16);
      }
    }
  }
  return AEMH_OK;
}

aemh_status aemh_fixed_write_particles(aemh_fixed* s, uint64_t first, uint64_t count, const int64_t* q29,
                                       const uint8_t* flags3, const uint64_t* ids9, const float* color4) {
  if (!s) return AEMH_ERR_ARG;
  FXCHK(check_records(s, first, count));
  if (count && (!q29 || !flags3 || !ids9 || !color4)) return fxfail(s, AEMH_ERR_ARG, "write_particles: null array");
  (void)hipSetDevice(s->device);
  const uint64_t chunk = 1u << 20;
  std::vector<Rec> recs;
  for (uint64_t b = 0; b < count; b += chunk) {
    const uint64_t m = std::min(chunk, count - b);
    recs.resize((size_t)m);
    for (uint64_t q = 0; q < m; ++q) {
      Rec& r = recs[(size_t)q];
      const uint64_t i = b + q;
      for (int k = 0; k < kFxScalarPlanes; ++k) r.w[k] = (uint64_t)q29[i * 29 + k];
      r.w[FX_FLAGS] = (uint64_t)(flags3[i * 3] != 0) | ((uint64_t)(flags3[i * 3 + 1] != 0) << 8) |
                      ((uint64_t)flags3[i * 3 + 2] << 16);
      for (int k = 0; k < 9; ++k) r.w[FX_ID + k] = ids9[i * 9 + k];
      r.w[FX_COLOR] = pack_color(color4[i * 4], color4[i * 4 + 1]);
      r.w[FX_COLOR + 1] = pack_color(color4[i * 4 + 2], color4[i * 4 + 3]);
    }
    FXCHK(put_records(s, (uint32_t)(first + b), recs.data(), (uint32_t)m));
  }
  return AEMH_OK;
}

aemh_status aemh_fixed_fracture_events(aemh_fixed* s, uint64_t first, uint64_t count, aemh_fracture_event* out) {
  if (!s || (count && !out)) return AEMH_ERR_ARG;
  if (first > s->n_events || count > s->n_events - first)
    return fxfail(s, AEMH_ERR_ARG, "events [%llu, +%llu) outside %u", (unsigned long long)first,
                  (unsigned long long)count, s->n_events);
  if (!count) return AEMH_OK;
  static_assert(sizeof(aemh_fracture_event) == sizeof(FxEvent), "event layout");
  (void)hipSetDevice(s->device);
  FXTRY(s, hipMemcpy(out, s->events + first, (size_t)count * sizeof(FxEvent), hipMemcpyDeviceToHost));
  return AEMH_OK;
}

aemh_status aemh_fixed_get_stats(aemh_fixed* s, aemh_fixed_stats* o) {
  if (!s || !o) return AEMH_ERR_ARG;
  (void)hipSetDevice(s->device);
  FxStatsAcc a;
  FXTRY(s, hipMemsetAsync(s->acc, 0, sizeof(FxStatsAcc), s->stream));
  FXTRY(s, fx_launch_stats(s->st, s->n, s->acc, s->stream));
  FXTRY(s, hipMemcpyAsync(&a, s->acc, sizeof(a), hipMemcpyDeviceToHost, s->stream));
  FXTRY(s, hipStreamSynchronize(s->stream));
  bool serial = a.tripped != 0;
  for (int k = 0; k < 4; ++k) serial |= a.bound[k] >= kFxBoundLimit;
  if (serial) {  // a saturating fold could clip: evaluate it in order
    FXTRY(s, fx_launch_stats_serial(s->st, s->n, s->acc, s->stream));
    FXTRY(s, hipMemcpyAsync(&a, s->acc, sizeof(a), hipMemcpyDeviceToHost, s->stream));
    FXTRY(s, hipStreamSynchronize(s->stream));
  }
  o->active_particles = a.active;
  o->fractured_count = a.fractured;
  o->total_mass = (int64_t)a.sum[0];
  o->kinetic_energy = (int64_t)a.sum[1];
  o->total_strain_energy = (int64_t)a.sum[2];
  // average_temperature (:756-764): sum / from_i32(count as i32)
  o->average_temperature = a.active == 0 ? from_f32(293.0f)
                                         : q32::div((int64_t)a.sum[3], from_int((int32_t)(uint32_t)a.active));
  return AEMH_OK;
}

aemh_status aemh_fixed_snapshots(aemh_fixed* s, aemh_fixed_snapshot* out, uint64_t cap, uint64_t* n_out) {
  if (!s || (cap && !out)) return AEMH_ERR_ARG;
  (void)hipSetDevice(s->device);
  uint64_t got = 0;
  if (s->n) {
    std::vector<FxSnapshot> all(s->n);
    FxSnapshot* d = nullptr;
    FXTRY(s, hipMalloc((void**)&d, (size_t)s->n * sizeof(FxSnapshot)));
    hipError_t e = fx_launch_snapshots(s->st, s->n, d, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(all.data(), d, all.size() * sizeof(FxSnapshot), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(d);
    FXTRY(s, e);
    for (const FxSnapshot& r : all) {
      if (!(r.fractured & 2)) continue;  // bit 1: active
      if (got < cap) {
        aemh_fixed_snapshot& o = out[got];
        memcpy(o.position, r.position, 12);
        memcpy(o.velocity, r.velocity, 12);
        o.stress_magnitude = r.stress_magnitude;
        o.plastic_strain = r.plastic_strain;
        o.temperature = r.temperature;
        memcpy(o.color, r.color, 16);
        o.fractured = r.fractured & 1;
      }
      ++got;
    }
  }
  if (n_out) *n_out = got;
  return AEMH_OK;
}

}  // extern "C"
