// k_fixed.h -- device data of the all-integer MPM solver (MpmSimulation,
// ae_physics/src/mpm.rs) and its launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// The particle store is a slab of 8-byte word planes, plane k of particle i
// at base[k * stride + i].  Planes 0..28 are the int64 fields in the order
// position(3), velocity(3), deformation gradient(9, column-major),
// stress(9, column-major), mass, volume, plastic strain, temperature,
// accumulated strain; then flags (byte 0 active, 1 fractured, 2 number of
// connections), parent id, connected ids(8), color (r|g, b|a as f32 pairs).
enum FxPlane {
  FX_POS = 0,
  FX_VEL = 3,
  FX_F = 6,
  FX_STRESS = 15,
  FX_MASS = 24,
  FX_VOLUME = 25,
  FX_PLASTIC = 26,
  FX_TEMP = 27,
  FX_ACCUM = 28,
  FX_FLAGS = 29,
  FX_ID = 30,
  FX_CONN = 31,
  FX_COLOR = 39,
  FX_NPLANES = 41
};
constexpr int kFxScalarPlanes = 29;

enum FxModel { FXM_ELASTIC = 0, FXM_ELASTOPLASTIC = 1, FXM_GRANULAR = 2, FXM_BRITTLE = 3, FXM_FLUID = 4, FXM_SNOW = 5 };

struct FxStore {
  int64_t* base = nullptr;
  size_t stride = 0;  // words between planes
};

// Constants of one substep, all computed on the host with the same rules.
struct FxSubstep {
  int64_t cell, inv_cell, sdt;
  int64_t gdt[3];          // gravity * dt
  int64_t mu, lambda, E, yield, hardening;
  int64_t mu_s;            // granular friction coefficient
  int64_t viscosity;       // fluid
  int64_t j_lo, j_hi;      // J clamp: from_f32(0.01), from_f32(100.0)
  int64_t crit_c, crit_s;  // snow singular-value clamps
  int32_t res[3];
  uint32_t n;
};

struct FxEvent {  // FractureEvent (mpm.rs:132-138)
  double time;
  float position[3];
  float strain_magnitude;
  uint32_t num_fragments;
  uint32_t pad;
};

struct FxSnapshot {  // MpmParticleSnapshot (mpm.rs:94-103)
  float position[3];
  float velocity[3];
  float stress_magnitude;
  float plastic_strain;
  float temperature;
  float color[4];
  uint32_t fractured;
};

// Accumulators of the order-free sums of the statistics: sums of mass,
// kinetic energy, strain and temperature, their bounds, the tripped flag,
// active and fractured counts.  (The P2G bound is four words: this
// substep's, tripped, the maximum over finished substeps, spare.)
struct FxStatsAcc {
  unsigned long long sum[4];
  unsigned long long bound[4];
  unsigned long long tripped;
  unsigned long long active, fractured;
};

constexpr uint64_t kFxBoundLimit = (uint64_t)1 << 39;  // sum of |c| / 2^24 below this: no saturation
constexpr int kFxFragTable = 9;                        // fragments per event: 1..9

// P2G with its exactness check: bound[4] (this substep's bound, its
// tripped flag, the maximum over finished substeps, any substep tripped;
// the grid kernel folds and clears the first two), cell_bound[ncells]
// zeroed, counters[2] (substeps checked per cell, rebuilt serially)
// The order P2G visits particles in: by base cell on the (res + 2)^3
// lattice (nbins = that size + 1, the last bin holding particles that
// scatter nothing; below 2^32).  P2G reads packed records in that order:
// position, velocity, stress, mass (kFxRecWords int64), written by the
// sort and then by each G2P at rank[i] together with the particle's new
// key.  Records, keys and perm (slot -> particle) are double-buffered; cur
// is the live copy.
constexpr int kFxRecWords = 16;
struct FxSortBufs {
  uint32_t* keys[2];
  int64_t* rec[2];
  uint32_t* perm[2];
  uint32_t* rank;    // particle -> slot
  uint32_t* bins;    // nbins counts
  uint32_t* cursor;  // nbins running offsets
  uint32_t nbins;
  int cur;
  void* temp;
  size_t temp_bytes;
};
enum { FX_ORDER_KEEP = 0, FX_ORDER_RESORT = 1, FX_ORDER_FROM_PLANES = 2 };
struct FxOrder {  // what G2P needs to refresh the records (rec == nullptr: nothing)
  const uint32_t* rank;
  int64_t* rec;
  uint32_t* keys;
  uint32_t none;
};
size_t fx_sort_temp_bytes(uint32_t nbins);
// sort == nullptr scatters in particle order with one atomic per value;
// else in the sort's order: kept (FX_ORDER_KEEP), re-sorted from the live
// records (FX_ORDER_RESORT), or rebuilt from the particle planes
// (FX_ORDER_FROM_PLANES: required after the particle set changed)
hipError_t fx_launch_p2g(const FxStore& st, const FxSubstep& sp, int64_t* grid,
                         unsigned long long* bound, unsigned long long* cell_bound, uint32_t ncells,
                         unsigned long long* counters, FxSortBufs* sort, int resort, hipStream_t s);
hipError_t fx_launch_grid(int64_t* grid, uint32_t ncells, const int64_t gdt[3], unsigned long long* bound,
                          hipStream_t s);
hipError_t fx_launch_g2p(int model, const FxStore& st, const FxSubstep& sp, const int64_t* grid,
                         const int64_t* sin_tab, const FxSortBufs* sort, hipStream_t s);
hipError_t fx_launch_frac_mark(const FxStore& st, uint32_t n, int64_t threshold, uint64_t* key,
                               int64_t* strain, hipStream_t s);
size_t fx_scan_temp_bytes(uint32_t n);
hipError_t fx_launch_scan(const uint64_t* key, uint64_t* off, uint32_t n, void* temp, size_t temp_bytes,
                          uint64_t* total_host, hipStream_t s);
hipError_t fx_launch_frac_spawn(const FxStore& st, uint32_t n, const uint64_t* key, const uint64_t* off,
                                const int64_t* strain, const int64_t* frag_off, uint64_t next_id,
                                double time, FxEvent* events, hipStream_t s);
hipError_t fx_launch_put(const FxStore& st, uint32_t at, const uint64_t* src, uint32_t m, hipStream_t s);
hipError_t fx_launch_get(const FxStore& st, uint32_t first, uint32_t m, uint64_t* dst, hipStream_t s);
hipError_t fx_launch_find(const FxStore& st, uint32_t n, uint64_t id, uint32_t* idx, hipStream_t s);
hipError_t fx_launch_stats(const FxStore& st, uint32_t n, FxStatsAcc* acc, hipStream_t s);
hipError_t fx_launch_stats_serial(const FxStore& st, uint32_t n, FxStatsAcc* acc, hipStream_t s);
hipError_t fx_launch_snapshots(const FxStore& st, uint32_t n, FxSnapshot* out, hipStream_t s);
