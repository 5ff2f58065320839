// k_fixed.hip -- kernels of the all-integer MLS-MPM solver MpmSimulation
// (ae_physics/src/mpm.rs:244-716) on the Q32.32 FixedPoint of q32.h.
//
// One thread per particle; the reference's per-particle loops have no
// cross-particle dependency except the grid scatter and fracture's appends:
//   * P2G adds int64 contributions with atomics.  Integer addition is
//     order-free unless a saturating prefix sum in the reference's serial
//     order would clip, and none can when the sum of |contribution| into a
//     value stays below 2^63: each thread adds (|c| >> 24) + 1 per nonzero
//     contribution to a whole-substep bound (limit 2^39).  Past it, a second
//     pass bounds each cell, and if a cell can reach the limit the grid is
//     rebuilt by the serial loop on the device.  The result is the
//     reference's in every case; the check costs two near-empty launches.
//   * Because the sums are order-free, large scenes scatter in an order of
//     their own: particles counting-sorted by base cell, read as packed
//     records that G2P refreshes, with the contributions of a wave's run of
//     equal cells summed by shuffles so one lane per run issues the atomic.
//   * grid forces and grid velocity (momentum / mass) are per cell; the
//     division is done once per cell instead of once per particle-node pair
//     in G2P (the same operands give the same quotient).
//   * G2P, update_deformation and the constitutive model are fused: each
//     is a loop over particles reading only that particle's fields.
//   * fracture marks candidates, an exclusive scan over (1 << 32 | nf)
//     gives each one its event index and first fragment slot, and the spawn
//     kernel writes fragments exactly where the serial appends put them.
#include <hipcub/hipcub.hpp>

#include "k_fixed.h"
#include "q32.h"

using namespace q32;

namespace {

constexpr int kFxBlock = 256;

__device__ inline float to_f32(int64_t r) { return (float)r / 4294967296.0f; }

// Rust `f32 as i32`: saturating, NaN -> 0
__device__ inline int32_t f32_to_i32(float f) {
  if (f != f) return 0;
  if (f >= 2147483648.0f) return INT32_MAX;
  if (f <= -2147483648.0f) return INT32_MIN;
  return (int32_t)f;
}

__device__ inline uint64_t bound_of(int64_t c) { return c == 0 ? 0 : (mag(c) >> 24) + 1; }

__device__ inline uint64_t wave_sum(uint64_t v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// adds a wave's bound to the substep total; a wave at the limit trips it
__device__ inline void push_bound(uint64_t b, unsigned long long* total, unsigned long long* tripped) {
  b = wave_sum(b);
  if ((threadIdx.x & 63) == 0 && b) {
    if (b >= kFxBoundLimit) {
      atomicOr(tripped, 1ull);
      b = kFxBoundLimit;
    }
    atomicAdd(total, (unsigned long long)b);
  }
}

struct PAt {
  const int64_t* base;
  size_t stride;
  uint32_t i;
  __device__ int64_t operator()(int plane) const { return base[(size_t)plane * stride + i]; }
};

// base cell (through f32, as the reference) and fractional offset per axis
__device__ inline void locate(int64_t pos, int64_t inv, int64_t& fx, int32_t& b) {
  const int64_t s = mul(pos, inv);
  const int64_t base = q32::floor(s);
  fx = sub(s, base);
  b = f32_to_i32(to_f32(base));
}

__device__ inline void weights(int64_t fx, int64_t w[3]) {
  for (int d = -1; d <= 1; ++d) w[d + 1] = bspline(sub(fx, from_int(d)));
}

// v[d + 1] for d in {-1, 0, 1} by selects (stays in registers in a loop)
__device__ inline int64_t sel3(const int64_t (&v)[3], int d) { return d < 0 ? v[0] : (d == 0 ? v[1] : v[2]); }

__device__ inline bool in_grid(int32_t x, int32_t y, int32_t z, const int32_t* res) {
  return x >= 0 && x < res[0] && y >= 0 && y < res[1] && z >= 0 && z < res[2];
}

__device__ inline size_t cell_of(int32_t x, int32_t y, int32_t z, const int32_t* res) {
  return (size_t)x + (size_t)y * (size_t)res[0] + (size_t)z * (size_t)res[0] * (size_t)res[1];
}

// The 27 node contributions of one particle in the reference's loop order
// (mpm.rs:286-333): fn(cell, mass*w, momentum*w [3]) for each node in range.
template <typename Fn>
__device__ inline void p2g_particle(const PAt& P, const FxSubstep& sp, Fn&& fn) {
  int64_t pos[3], vel[3], fx[3], w[3][3];
  int32_t b[3];
  Mat stress;
  for (int a = 0; a < 3; ++a) {
    pos[a] = P(FX_POS + a);
    vel[a] = P(FX_VEL + a);
    locate(pos[a], sp.inv_cell, fx[a], b[a]);
    weights(fx[a], w[a]);
  }
  for (int k = 0; k < 9; ++k) stress.m[k] = P(FX_STRESS + k);
  const int64_t mass = P(FX_MASS);
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        const int32_t gx = b[0] + dx, gy = b[1] + dy, gz = b[2] + dz;
        if (!in_grid(gx, gy, gz, sp.res)) continue;
        const int64_t wt = mul(mul(w[0][dx + 1], w[1][dy + 1]), w[2][dz + 1]);
        if (wt < EPS) continue;
        const int64_t rel[3] = {sub(mul(from_int(gx), sp.cell), pos[0]), sub(mul(from_int(gy), sp.cell), pos[1]),
                                sub(mul(from_int(gz), sp.cell), pos[2])};
        int64_t sr[3];
        matvec(stress, rel, sr);
        int64_t v[4];
        v[0] = mul(mass, wt);
        for (int a = 0; a < 3; ++a) v[1 + a] = mul(mul(mass, add(vel[a], mul(mul(sr[a], sp.inv_cell), sp.inv_cell))), wt);
        fn(cell_of(gx, gy, gz, sp.res), v);
      }
}

__device__ inline bool bound_ok(const unsigned long long* bound) {
  return bound[0] < kFxBoundLimit && bound[1] == 0;
}

// particles_to_grid, order-free: int64 atomics, plus the bound on the sum of
// |contribution| that shows no saturating prefix sum could have clipped
__global__ __launch_bounds__(kFxBlock) void k_fx_p2g(FxStore st, FxSubstep sp, int64_t* __restrict__ grid,
                                                      unsigned long long* __restrict__ bound) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t bsum = 0;
  if (i < sp.n) {
    const PAt P{st.base, st.stride, i};
    if (P(FX_FLAGS) & 1)
      p2g_particle(P, sp, [&](size_t c, const int64_t* v) {
        for (int k = 0; k < 4; ++k)
          if (v[k] != 0) {
            atomicAdd(reinterpret_cast<unsigned long long*>(grid + c * 4 + k), (unsigned long long)v[k]);
            bsum += bound_of(v[k]);
          }
      });
  }
  push_bound(bsum, bound, bound + 1);
}

// Aggregation key of a particle: its base cell on the (res + 2)^3 lattice
// that also holds the bases one cell outside the grid (their +1 nodes are
// inside); `none` for inactive particles and bases further out, which
// scatter nothing.
__device__ inline uint32_t agg_key(const PAt& P, const FxSubstep& sp, uint32_t none) {
  if (!(P(FX_FLAGS) & 1)) return none;
  int32_t b[3];
  for (int a = 0; a < 3; ++a) {
    int64_t fx;
    locate(P(FX_POS + a), sp.inv_cell, fx, b[a]);
    if (b[a] < -1 || b[a] > sp.res[a]) return none;
  }
  const uint32_t ex = (uint32_t)sp.res[0] + 2, ey = (uint32_t)sp.res[1] + 2;
  return (uint32_t)(b[0] + 1) + ex * ((uint32_t)(b[1] + 1) + ey * (uint32_t)(b[2] + 1));
}

// Counting sort of the P2G order by base cell (k_fx_bin_count, a scan of
// the bins, k_fx_bin_scatter).  Not stable: particles of one cell land in
// any order, which the exact integer scatter does not see.  The source is
// either the particle planes in particle order (after the particle set
// changed), keys computed here, or the previous order's records and keys.

// Runs of equal keys within a wave (a sorted source has long ones): the
// run's first lane and each lane's position in its run.  Every lane calls
// it (lanes past the end pass a key of their own that matches no other).
struct WaveRun {
  bool head;
  uint32_t pos, len;
};
__device__ inline WaveRun wave_run(uint32_t key) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t prev = __shfl_up(key, 1);
  const bool head = lane == 0 || prev != key;
  uint32_t rid = head ? lane : 0u;  // lane of the run's head
  for (int k = 0; k < 6; ++k) {
    const uint32_t o = __shfl_up(rid, 1u << k);
    if (lane >= (1u << k) && o > rid) rid = o;
  }
  // the run ends where the next head starts: lanes of the run count up to it
  const uint64_t heads = __ballot(head);
  const uint64_t later = heads & ~((2ull << lane) - 1);  // heads after this lane
  const uint32_t end = later ? (uint32_t)__builtin_ctzll(later) : 64u;
  return WaveRun{head, lane - rid, end - rid};
}

__global__ __launch_bounds__(kFxBlock) void k_fx_bin_count(FxStore st, FxSubstep sp, uint32_t none, bool from_planes,
                                                            uint32_t* __restrict__ keys, uint32_t* __restrict__ bins) {
  const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = u < sp.n;
  uint32_t key = 0xffffffffu - (threadIdx.x & 63);  // distinct per idle lane
  if (in) {
    if (from_planes) {
      key = agg_key(PAt{st.base, st.stride, u}, sp, none);
      keys[u] = key;
    } else {
      key = keys[u];
    }
  }
  const WaveRun r = wave_run(key);
  if (in && r.head) atomicAdd(bins + key, r.len);
}

__global__ __launch_bounds__(kFxBlock) void k_fx_bin_scatter(FxStore st, uint32_t n, bool from_planes,
                                                              const uint32_t* __restrict__ keys,
                                                              const int64_t* __restrict__ rec,
                                                              const uint32_t* __restrict__ perm,
                                                              uint32_t* __restrict__ cursor,
                                                              uint32_t* __restrict__ keys_out,
                                                              int64_t* __restrict__ rec_out,
                                                              uint32_t* __restrict__ perm_out,
                                                              uint32_t* __restrict__ rank) {
  const uint32_t u = blockIdx.x * blockDim.x + threadIdx.x;
  const bool in = u < n;
  const uint32_t key = in ? keys[u] : 0xffffffffu - (threadIdx.x & 63);
  const WaveRun run = wave_run(key);
  uint32_t base = 0;
  if (in && run.head) base = atomicAdd(cursor + key, run.len);
  // the head's slot, from the lane that holds it
  base = __shfl(base, (int)((threadIdx.x & 63) - run.pos));
  if (!in) return;
  const uint32_t v = base + run.pos;
  int64_t* r = rec_out + (size_t)v * kFxRecWords;
  uint32_t i;
  if (from_planes) {
    i = u;
    const PAt P{st.base, st.stride, u};
    for (int k = 0; k < 6; ++k) r[k] = P(FX_POS + k);  // position, velocity
    for (int k = 0; k < 9; ++k) r[6 + k] = P(FX_STRESS + k);
    r[15] = P(FX_MASS);
  } else {
    i = perm[u];
    const int64_t* q = rec + (size_t)u * kFxRecWords;
    for (int k = 0; k < kFxRecWords; ++k) r[k] = q[k];
  }
  keys_out[v] = key;
  perm_out[v] = i;
  rank[i] = v;
}

// particles_to_grid over the particles in (roughly) key order, reading
// their packed records: lanes of a wave in one run of equal base cells
// share all 27 nodes, so each node value is summed over the run (a
// segmented reduction by shuffles) and only the run's first lane adds it to
// the grid.  Runs are maximal stretches of equal adjacent keys, named by
// their first lane, so an order left over from an earlier sort (equal keys
// split into several runs) is still summed exactly.  The sums are wrapping
// int64 like the atomics, and the bound still counts every particle's own
// contributions.
__global__ __launch_bounds__(kFxBlock) void k_fx_p2g_sorted(FxSubstep sp, const uint32_t* __restrict__ skeys,
                                                             const int64_t* __restrict__ rec, uint32_t none,
                                                             int64_t* __restrict__ grid,
                                                             unsigned long long* __restrict__ bound) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t key = t < sp.n ? skeys[t] : none;
  // the shuffles run on every lane: a lane masked off (by a short-circuit,
  // say) would hand its neighbour an undefined value
  const uint32_t prev = __shfl_up(key, 1);
  const bool head = lane == 0 || prev != key;
  // run id: the lane of the run's head (inclusive max-scan of head lanes)
  uint32_t rid = head ? lane : 0u;
  for (int k = 0; k < 6; ++k) {
    const uint32_t o = __shfl_up(rid, 1u << k);
    if (lane >= (1u << k) && o > rid) rid = o;
  }
  // same bit k: lane + 2^k is in this lane's run (so is every lane between)
  uint32_t same = 0;
  for (int k = 0; k < 6; ++k) {
    const uint32_t o = 1u << k;
    const uint32_t other = __shfl_down(rid, o);
    if (lane + o < 64 && other == rid) same |= 1u << k;
  }
  const bool live = key != none;
  int64_t pos[3] = {0, 0, 0}, vel[3] = {0, 0, 0}, fx[3] = {0, 0, 0}, w[3][3];
  int32_t b[3] = {0, 0, 0};
  Mat stress = mat_zero();
  int64_t mass = 0;
  if (live) {
    const int64_t* r = rec + (size_t)t * kFxRecWords;
    for (int a = 0; a < 3; ++a) {
      pos[a] = r[a];
      vel[a] = r[3 + a];
      locate(pos[a], sp.inv_cell, fx[a], b[a]);
      weights(fx[a], w[a]);
    }
    for (int k = 0; k < 9; ++k) stress.m[k] = r[6 + k];
    mass = r[15];
  }
  // Component a of node (dx, dy, dz) minus the position depends on one
  // axis offset only, and so does each product of the stress-times-offset
  // sums (row r, column a): term[a][r][d] = stress(r, a) * rel_a(d).
  int64_t term[3][3][3];
  for (int a = 0; a < 3; ++a)
    for (int d = 0; d < 3; ++d) {
      const int64_t rel = sub(mul(from_int(b[a] + d - 1), sp.cell), pos[a]);
      for (int r = 0; r < 3; ++r) term[a][r][d] = mul(stress.at(r, a), rel);
    }
  // every lane of a run that reaches a head has the same base; a wave with
  // no repeated key skips the reductions
  const bool any_run = __any(same != 0);
  uint64_t bsum = 0;
  for (int dx = -1; dx <= 1; ++dx) {
    const int64_t wx = sel3(w[0], dx);
    int64_t tx[3];
    for (int r = 0; r < 3; ++r) tx[r] = sel3(term[0][r], dx);
    for (int dy = -1; dy <= 1; ++dy) {
      const int64_t wy = sel3(w[1], dy);
      int64_t ty[3];
      for (int r = 0; r < 3; ++r) ty[r] = sel3(term[1][r], dy);
      for (int dz = -1; dz <= 1; ++dz) {
        const int32_t gx = b[0] + dx, gy = b[1] + dy, gz = b[2] + dz;
        const bool node = live && in_grid(gx, gy, gz, sp.res);
        int64_t v[4] = {0, 0, 0, 0};
        if (node) {
          const int64_t wt = mul_w(mul_w(wx, (uint32_t)wy), (uint32_t)sel3(w[2], dz));
          if (wt >= EPS) {
            int64_t sr[3];  // stress * rel, each row summed x, y, z (matvec)
            for (int r = 0; r < 3; ++r) sr[r] = add(add(tx[r], ty[r]), sel3(term[2][r], dz));
            v[0] = mul_w(mass, (uint32_t)wt);
            for (int a = 0; a < 3; ++a)
              v[1 + a] = mul_w(mul(mass, add(vel[a], mul(mul(sr[a], sp.inv_cell), sp.inv_cell))), (uint32_t)wt);
            for (int k = 0; k < 4; ++k) bsum += bound_of(v[k]);
          }
        }
        if (any_run) {
          for (int k = 0; k < 4; ++k) {
            uint64_t x = (uint64_t)v[k];
            for (int q = 0; q < 6; ++q) {
              const uint64_t o = __shfl_down(x, 1u << q);
              if (same & (1u << q)) x += o;
            }
            v[k] = (int64_t)x;
          }
        }
        if (node && head) {
          int64_t* g = grid + cell_of(gx, gy, gz, sp.res) * 4;
          for (int k = 0; k < 4; ++k)
            if (v[k] != 0) atomicAdd(reinterpret_cast<unsigned long long*>(g + k), (unsigned long long)v[k]);
        }
      }
    }
  }
  push_bound(bsum, bound, bound + 1);
}

// when the whole-substep bound is exceeded: the same bound per cell
__global__ __launch_bounds__(kFxBlock) void k_fx_cell_bound(FxStore st, FxSubstep sp,
                                                             const unsigned long long* __restrict__ bound,
                                                             unsigned long long* __restrict__ cb) {
  if (bound_ok(bound)) return;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sp.n) return;
  const PAt P{st.base, st.stride, i};
  if (!(P(FX_FLAGS) & 1)) return;
  p2g_particle(P, sp, [&](size_t c, const int64_t* v) {
    uint64_t b = 0;
    for (int k = 0; k < 4; ++k) b += bound_of(v[k]);
    if (b) atomicAdd(cb + c, (unsigned long long)b);
  });
}

// Decides whether the atomic scatter is exact: nothing to do when the
// whole-substep bound holds; else the largest per-cell bound (cleared on the
// way); and if some cell could have saturated, the grid is rebuilt by the
// reference's serial loop, one thread per accumulated value (mass, momentum
// x, y, z are independent sums), with saturating adds in particle order.
// counters[0]: substeps that needed the per-cell check, [1]: serial rebuilds.
__global__ __launch_bounds__(256) void k_fx_p2g_settle(FxStore st, FxSubstep sp, int64_t* __restrict__ grid,
                                                       unsigned long long* __restrict__ bound,
                                                       unsigned long long* __restrict__ cb, uint32_t ncells,
                                                       unsigned long long* __restrict__ counters) {
  if (bound_ok(bound)) return;
  __shared__ unsigned long long wmax[4];
  unsigned long long m = 0;
  for (uint32_t c = threadIdx.x; c < ncells; c += blockDim.x) {
    const unsigned long long v = cb[c];
    if (v > m) m = v;
    if (v) cb[c] = 0;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long t = __shfl_xor(m, o);
    if (t > m) m = t;
  }
  if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
  __syncthreads();
  m = wmax[0];
  for (int k = 1; k < (int)(blockDim.x >> 6); ++k)
    if (wmax[k] > m) m = wmax[k];
  const bool serial = m >= kFxBoundLimit;
  if (serial) {
    for (size_t k = threadIdx.x; k < (size_t)ncells * 4; k += blockDim.x) grid[k] = 0;
    __syncthreads();
    if (threadIdx.x < 4) {
      const int k = threadIdx.x;
      for (uint32_t i = 0; i < sp.n; ++i) {
        const PAt P{st.base, st.stride, i};
        if (!(P(FX_FLAGS) & 1)) continue;
        p2g_particle(P, sp, [&](size_t c, const int64_t* v) { grid[c * 4 + k] = add(grid[c * 4 + k], v[k]); });
      }
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) counters[serial ? 1 : 0] += 1;
}

// compute_grid_forces (mpm.rs:335-344) and the grid velocity G2P divides by;
// also folds the finished P2G's bound into the maximum over substeps and
// clears it for the next one
__global__ __launch_bounds__(kFxBlock) void k_fx_grid(int64_t* __restrict__ grid, uint32_t ncells, int64_t g0,
                                                       int64_t g1, int64_t g2, unsigned long long* bound) {
  const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c == 0) {
    if (bound[0] > bound[2]) bound[2] = bound[0];
    if (bound[1]) bound[3] = 1;
    bound[0] = 0;
    bound[1] = 0;
  }
  if (c >= ncells) return;
  int64_t* g = grid + (size_t)c * 4;
  const int64_t m = g[0];
  if (m < EPS) return;
  g[1] = q32::div(add(g[1], mul(g0, m)), m);
  g[2] = q32::div(add(g[2], mul(g1, m)), m);
  g[3] = q32::div(add(g[3], mul(g2, m)), m);
}

__device__ inline Mat plastic_return(const Mat& stress, int64_t h, const Mat& dev, int64_t scale) {
  return mat_add(mat_diag(h), mat_scale(dev, scale));
}

// grid_to_particles (mpm.rs:346-427), update_deformation (:429-453) and the
// constitutive model (:456-639) of one particle
template <int MODEL>
__global__ __launch_bounds__(kFxBlock) void k_fx_g2p(FxStore st, FxSubstep sp, const int64_t* __restrict__ grid,
                                                      const int64_t* __restrict__ tab, FxOrder ord) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= sp.n) return;
  const PAt P{st.base, st.stride, i};
  if (!(P(FX_FLAGS) & 1)) return;
  int64_t pos[3], fx[3], w[3][3];
  int32_t b[3];
  for (int a = 0; a < 3; ++a) {
    pos[a] = P(FX_POS + a);
    locate(pos[a], sp.inv_cell, fx[a], b[a]);
    weights(fx[a], w[a]);
  }
  int64_t nv[3] = {0, 0, 0};
  Mat vg = mat_zero();
  for (int dx = -1; dx <= 1; ++dx)
    for (int dy = -1; dy <= 1; ++dy)
      for (int dz = -1; dz <= 1; ++dz) {
        const int32_t gx = b[0] + dx, gy = b[1] + dy, gz = b[2] + dz;
        if (!in_grid(gx, gy, gz, sp.res)) continue;
        const int64_t wt = mul_w(w[0][dx + 1], (uint32_t)w[1][dy + 1]);
        const int64_t wt3 = mul_w(wt, (uint32_t)w[2][dz + 1]);
        if (wt3 < EPS) continue;
        const int64_t* g = grid + cell_of(gx, gy, gz, sp.res) * 4;
        if (g[0] < EPS) continue;
        const int64_t gv[3] = {g[1], g[2], g[3]};
        const int64_t gd[3] = {sub(from_int(dx), fx[0]), sub(from_int(dy), fx[1]), sub(from_int(dz), fx[2])};
        for (int a = 0; a < 3; ++a) {
          const int64_t gw = mul_w(gv[a], (uint32_t)wt3);  // grid_vel * weight
          nv[a] = add(nv[a], gw);
          for (int c = 0; c < 3; ++c) vg.m[a * 3 + c] = add(vg.m[a * 3 + c], mul(mul(gw, gd[c]), sp.inv_cell));
        }
      }
  int64_t* out = st.base + i;
  const size_t S = st.stride;
  int64_t vel[3];
  for (int a = 0; a < 3; ++a) {
    vel[a] = nv[a];
    out[(FX_POS + a) * S] = add(pos[a], mul(nv[a], sp.sdt));
  }
  Mat F;
  for (int k = 0; k < 9; ++k) F.m[k] = P(FX_F + k);
  F = mat_add(F, mat_scale(matmul(vg, F), sp.sdt));
  {
    const int64_t dv = det(vg);
    const int64_t strain = dv > ONE ? sub(dv, ONE) : sub(ONE, dv);
    out[FX_ACCUM * S] = add(P(FX_ACCUM), strain);
  }
  int64_t plastic = P(FX_PLASTIC);
  Mat stress;
  if (MODEL != FXM_SNOW) {  // snow recomputes the stress from its clamped F
    const int64_t j = det(F);
    if (qabs(j) < EPS) {
      for (int k = 0; k < 9; ++k) stress.m[k] = P(FX_STRESS + k);
    } else {
      stress = kirchhoff(F, j, sp.mu, sp.lambda, sp.j_lo, sp.j_hi);
    }
  }
  if (MODEL == FXM_ELASTOPLASTIC || MODEL == FXM_BRITTLE) {  // :474-502
    const int64_t h = div_int(trace(stress), 3);
    const Mat dev = mat_sub(stress, mat_diag(h));
    const int64_t vm = q32::sqrt(mul(ONE_HALF, dev_norm(dev)));
    const int64_t limit = add(sp.yield, mul(sp.hardening, plastic));
    if (vm > limit) {
      stress = plastic_return(stress, h, dev, q32::div(limit, vm));
      plastic = add(plastic, q32::div(sub(vm, limit), sp.E));
    }
  } else if (MODEL == FXM_GRANULAR) {  // :504-540
    const int64_t tr = trace(stress);
    const int64_t pressure = div_int(neg(tr), 3);
    if (pressure < 0) {
      stress = mat_zero();
    } else {
      const int64_t h = div_int(tr, 3);
      const Mat dev = mat_sub(stress, mat_diag(h));
      const int64_t dn = dev_norm(dev);
      const int64_t max_shear = add(mul(sp.mu_s, pressure), sp.yield);
      if (dn > max_shear) {
        stress = plastic_return(stress, h, dev, q32::div(max_shear, dn));
        plastic = add(plastic, q32::div(sub(dn, max_shear), sp.E));
      }
    }
  } else if (MODEL == FXM_FLUID) {  // :542-562
    const int64_t pressure = div_int(neg(trace(stress)), 3);
    stress = pressure < 0 ? mat_zero() : mat_diag(-pressure);
    const int64_t vm = q32::sqrt(dot3(vel[0], vel[1], vel[2], vel[0], vel[1], vel[2]));
    if (vm > 0) {
      const int64_t damp = q32::exp(mul(neg(sp.viscosity), vm));
      for (int a = 0; a < 3; ++a) vel[a] = mul(vel[a], damp);
    }
  } else if (MODEL == FXM_SNOW) {  // :564-639
    Mat U, V;
    int64_t sv[3];
    svd(F, U, sv, V, tab);
    for (int a = 0; a < 3; ++a) {
      if (sv[a] < sp.crit_c)
        sv[a] = sp.crit_c;
      else if (sv[a] > sp.crit_s)
        sv[a] = sp.crit_s;
    }
    Mat S = mat_zero();
    S.m[0] = sv[0];
    S.m[4] = sv[1];
    S.m[8] = sv[2];
    F = matmul(matmul(U, S), transpose(V));
    stress = kirchhoff(F, det(F), sp.mu, sp.lambda, sp.j_lo, sp.j_hi);
    const int64_t h = div_int(trace(stress), 3);
    const Mat dev = mat_sub(stress, mat_diag(h));
    const int64_t vm = q32::sqrt(mul(ONE_HALF, dev_norm(dev)));
    if (vm > sp.hardening) {
      stress = plastic_return(stress, h, dev, q32::div(sp.hardening, vm));
      plastic = add(plastic, q32::div(sub(vm, sp.hardening), sp.E));
    }
  }
  for (int a = 0; a < 3; ++a) out[(FX_VEL + a) * S] = vel[a];
  for (int k = 0; k < 9; ++k) {
    out[(FX_F + k) * S] = F.m[k];
    out[(FX_STRESS + k) * S] = stress.m[k];
  }
  out[FX_PLASTIC * S] = plastic;
  if (ord.rec) {  // the next P2G's record and key, at this particle's slot in its order
    const uint32_t t = ord.rank[i];
    int64_t* r = ord.rec + (size_t)t * kFxRecWords;
    int32_t nb[3];
    bool inside = true;
    for (int a = 0; a < 3; ++a) {
      const int64_t np = add(pos[a], mul(nv[a], sp.sdt));
      r[a] = np;
      r[3 + a] = vel[a];
      int64_t f;
      locate(np, sp.inv_cell, f, nb[a]);
      inside &= nb[a] >= -1 && nb[a] <= sp.res[a];
    }
    for (int k = 0; k < 9; ++k) r[6 + k] = stress.m[k];
    r[15] = P(FX_MASS);
    const uint32_t ex = (uint32_t)sp.res[0] + 2, ey = (uint32_t)sp.res[1] + 2;
    ord.keys[t] = inside ? (uint32_t)(nb[0] + 1) + ex * ((uint32_t)(nb[1] + 1) + ey * (uint32_t)(nb[2] + 1)) : ord.none;
  }
}

// handle_fracture (mpm.rs:641-716), first pass: which particles break and
// into how many fragments
__global__ __launch_bounds__(kFxBlock) void k_fx_frac_mark(FxStore st, uint32_t n, int64_t thr,
                                                            uint64_t* __restrict__ key,
                                                            int64_t* __restrict__ strain_out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const PAt P{st.base, st.stride, i};
  const uint64_t flags = (uint64_t)P(FX_FLAGS);
  uint64_t k = 0;
  if ((flags & 1) && !((flags >> 8) & 1)) {
    int64_t tr = 0;
    int64_t diag[3];
    for (int c = 0; c < 3; ++c) {  // diagonal of F^T F
      const int64_t f0 = P(FX_F + c * 3), f1 = P(FX_F + c * 3 + 1), f2 = P(FX_F + c * 3 + 2);
      diag[c] = dot3(f0, f1, f2, f0, f1, f2);
    }
    tr = add(add(diag[0], diag[1]), diag[2]);
    const int64_t ps = div_int(tr, 3);
    const int64_t strain = ps > ONE ? sub(ps, ONE) : sub(ONE, ps);
    if (strain > thr) {
      const float rf = fminf(to_f32(q32::div(strain, thr)), 8.0f);
      const uint32_t nf = (rf > 0.0f ? (uint32_t)rf : 0u) + 1u;
      k = ((uint64_t)1 << 32) | nf;
      strain_out[i] = strain;
    }
  }
  key[i] = k;
}

__global__ void k_fx_total(const uint64_t* __restrict__ key, const uint64_t* __restrict__ off, uint32_t n,
                           uint64_t* total) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *total = n ? off[n - 1] + key[n - 1] : 0;
}

// second pass: fragments at the slots the serial appends give them
__global__ __launch_bounds__(kFxBlock) void k_fx_frac_spawn(FxStore st, uint32_t n, const uint64_t* __restrict__ key,
                                                             const uint64_t* __restrict__ off,
                                                             const int64_t* __restrict__ strain,
                                                             const int64_t* __restrict__ frag_off, uint64_t next_id,
                                                             double time, FxEvent* __restrict__ events) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint64_t k = key[i];
  if (!k) return;
  const uint32_t nf = (uint32_t)k;
  const uint64_t o = off[i];
  const uint32_t ev = (uint32_t)(o >> 32), f0 = (uint32_t)o;
  int64_t* B = st.base;
  const size_t S = st.stride;
  const int64_t pos[3] = {B[FX_POS * S + i], B[(FX_POS + 1) * S + i], B[(FX_POS + 2) * S + i]};
  const int64_t vel[3] = {B[FX_VEL * S + i], B[(FX_VEL + 1) * S + i], B[(FX_VEL + 2) * S + i]};
  const int64_t mass = B[FX_MASS * S + i], volume = B[FX_VOLUME * S + i];
  const uint64_t parent = (uint64_t)B[FX_ID * S + i];
  const int64_t col0 = B[FX_COLOR * S + i], col1 = B[(FX_COLOR + 1) * S + i];
  const int64_t div_by = (int64_t)nf + 1;  // num_fragments_fp = from_i32(nf + 1)
  const uint32_t ncon = nf < 7 ? nf : 7;
  const int64_t flags = 1 | ((int64_t)(nf < 8 ? nf : 8) << 16);
  for (uint32_t fi = 0; fi < nf; ++fi) {
    const size_t j = (size_t)n + f0 + fi;
    const int64_t* ofs = frag_off + ((size_t)(nf - 1) * kFxFragTable + fi) * 3;
    for (int a = 0; a < 3; ++a) {
      B[(FX_POS + a) * S + j] = add(pos[a], ofs[a]);
      B[(FX_VEL + a) * S + j] = add(vel[a], mul(ofs[a], TWO));
    }
    for (int q = 0; q < 9; ++q) {
      B[(FX_F + q) * S + j] = (q == 0 || q == 4 || q == 8) ? ONE : 0;
      B[(FX_STRESS + q) * S + j] = 0;
    }
    B[FX_MASS * S + j] = div_int(mass, div_by);
    B[FX_VOLUME * S + j] = div_int(volume, div_by);
    B[FX_PLASTIC * S + j] = 0;
    B[FX_TEMP * S + j] = (int64_t)293 << 32;  // from_f32(293.0)
    B[FX_ACCUM * S + j] = 0;
    B[FX_FLAGS * S + j] = flags;
    B[FX_ID * S + j] = (int64_t)(next_id + f0 + fi);
    for (uint32_t c = 0; c < 8; ++c) B[(FX_CONN + c) * S + j] = c < ncon ? (int64_t)(parent + c + 1) : 0;
    B[FX_COLOR * S + j] = col0;
    B[(FX_COLOR + 1) * S + j] = col1;
  }
  B[FX_MASS * S + i] = div_int(mass, div_by);
  B[FX_FLAGS * S + i] |= (int64_t)1 << 8;
  FxEvent e;
  e.time = time;
  for (int a = 0; a < 3; ++a) e.position[a] = to_f32(pos[a]);
  e.strain_magnitude = to_f32(strain[i]);
  e.num_fragments = nf;
  e.pad = 0;
  events[ev] = e;
}

// AoS records of FX_NPLANES words <-> the plane store
__global__ void k_fx_put(FxStore st, uint32_t at, const uint64_t* __restrict__ src, uint32_t m) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)m * FX_NPLANES) return;
  const size_t p = t / FX_NPLANES, k = t % FX_NPLANES;
  st.base[k * st.stride + at + p] = (int64_t)src[t];
}
__global__ void k_fx_get(FxStore st, uint32_t first, uint32_t m, uint64_t* __restrict__ dst) {
  const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (size_t)m * FX_NPLANES) return;
  const size_t p = t / FX_NPLANES, k = t % FX_NPLANES;
  dst[t] = (uint64_t)st.base[k * st.stride + first + p];
}

__global__ void k_fx_find(FxStore st, uint32_t n, uint64_t id, uint32_t* idx) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && (uint64_t)st.base[FX_ID * st.stride + i] == id) atomicMin(idx, i);
}

// total_mass, kinetic_energy, total_strain_energy and the temperature sum
// of average_temperature (mpm.rs:744-778): order-free when no saturation
// can occur, else the serial kernel below
__global__ __launch_bounds__(kFxBlock) void k_fx_stats(FxStore st, uint32_t n, FxStatsAcc* acc) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  int64_t t[4] = {0, 0, 0, 0};
  uint64_t act = 0, frac = 0;
  if (i < n) {
    const PAt P{st.base, st.stride, i};
    const uint64_t flags = (uint64_t)P(FX_FLAGS);
    frac = (flags >> 8) & 1;
    if (flags & 1) {
      act = 1;
      const int64_t m = P(FX_MASS), vx = P(FX_VEL), vy = P(FX_VEL + 1), vz = P(FX_VEL + 2);
      t[0] = m;
      t[1] = mul(mul(HALF, m), dot3(vx, vy, vz, vx, vy, vz));
      t[2] = P(FX_ACCUM);
      t[3] = P(FX_TEMP);
    }
  }
  for (int k = 0; k < 4; ++k) {
    const uint64_t s = wave_sum((uint64_t)t[k]);
    if ((threadIdx.x & 63) == 0 && s) atomicAdd(&acc->sum[k], (unsigned long long)s);
    push_bound(bound_of(t[k]), &acc->bound[k], &acc->tripped);
  }
  act = wave_sum(act);
  frac = wave_sum(frac);
  if ((threadIdx.x & 63) == 0) {
    if (act) atomicAdd(&acc->active, (unsigned long long)act);
    if (frac) atomicAdd(&acc->fractured, (unsigned long long)frac);
  }
}

__global__ void k_fx_stats_serial(FxStore st, uint32_t n, FxStatsAcc* acc) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int64_t s[4] = {0, 0, 0, 0};
  for (uint32_t i = 0; i < n; ++i) {
    const PAt P{st.base, st.stride, i};
    if (!(P(FX_FLAGS) & 1)) continue;
    const int64_t m = P(FX_MASS), vx = P(FX_VEL), vy = P(FX_VEL + 1), vz = P(FX_VEL + 2);
    s[0] = add(s[0], m);
    s[1] = add(s[1], mul(mul(HALF, m), dot3(vx, vy, vz, vx, vy, vz)));
    s[2] = add(s[2], P(FX_ACCUM));
    s[3] = add(s[3], P(FX_TEMP));
  }
  for (int k = 0; k < 4; ++k) acc->sum[k] = (unsigned long long)s[k];
}

// snapshots (mpm.rs:718-742) of every particle; the host keeps active ones
__global__ __launch_bounds__(kFxBlock) void k_fx_snap(FxStore st, uint32_t n, FxSnapshot* __restrict__ out) {
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const PAt P{st.base, st.stride, i};
  FxSnapshot r;
  Mat s;
  for (int k = 0; k < 9; ++k) s.m[k] = P(FX_STRESS + k);
  for (int a = 0; a < 3; ++a) {
    r.position[a] = to_f32(P(FX_POS + a));
    r.velocity[a] = to_f32(P(FX_VEL + a));
  }
  r.stress_magnitude = to_f32(dev_norm(s));
  r.plastic_strain = to_f32(P(FX_PLASTIC));
  r.temperature = to_f32(P(FX_TEMP));
  const uint64_t c0 = (uint64_t)P(FX_COLOR), c1 = (uint64_t)P(FX_COLOR + 1);
  r.color[0] = __uint_as_float((uint32_t)c0);
  r.color[1] = __uint_as_float((uint32_t)(c0 >> 32));
  r.color[2] = __uint_as_float((uint32_t)c1);
  r.color[3] = __uint_as_float((uint32_t)(c1 >> 32));
  r.fractured = (uint32_t)(((uint64_t)P(FX_FLAGS) >> 8) & 1) | ((uint32_t)((uint64_t)P(FX_FLAGS) & 1) << 1);
  out[i] = r;
}

inline unsigned blocks_for(size_t n) { return (unsigned)((n + kFxBlock - 1) / kFxBlock); }

}  // namespace

size_t fx_sort_temp_bytes(uint32_t nbins) {
  size_t bytes = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const uint32_t*)nullptr, (uint32_t*)nullptr, (int)nbins);
  return bytes;
}

hipError_t fx_launch_p2g(const FxStore& st, const FxSubstep& sp, int64_t* grid, unsigned long long* bound,
                         unsigned long long* cell_bound, uint32_t ncells, unsigned long long* counters,
                         FxSortBufs* sort, int resort, hipStream_t s) {
  if (!sp.n) return hipSuccess;
  if (sort) {
    const uint32_t none = sort->nbins - 1;
    if (resort != FX_ORDER_KEEP) {
      const bool planes = resort == FX_ORDER_FROM_PLANES;
      const int src = planes ? 1 - sort->cur : sort->cur, dst = 1 - src;
      hipError_t e = hipMemsetAsync(sort->bins, 0, (size_t)sort->nbins * 4, s);
      if (e != hipSuccess) return e;
      k_fx_bin_count<<<blocks_for(sp.n), kFxBlock, 0, s>>>(st, sp, none, planes, sort->keys[src], sort->bins);
      size_t bytes = sort->temp_bytes;
      e = hipcub::DeviceScan::ExclusiveSum(sort->temp, bytes, sort->bins, sort->cursor, (int)sort->nbins, s);
      if (e != hipSuccess) return e;
      k_fx_bin_scatter<<<blocks_for(sp.n), kFxBlock, 0, s>>>(st, sp.n, planes, sort->keys[src], sort->rec[src],
                                                              sort->perm[src], sort->cursor, sort->keys[dst],
                                                              sort->rec[dst], sort->perm[dst], sort->rank);
      sort->cur = dst;
    }
    k_fx_p2g_sorted<<<blocks_for(sp.n), kFxBlock, 0, s>>>(sp, sort->keys[sort->cur], sort->rec[sort->cur], none,
                                                            grid, bound);
  } else {
    k_fx_p2g<<<blocks_for(sp.n), kFxBlock, 0, s>>>(st, sp, grid, bound);
  }
  k_fx_cell_bound<<<blocks_for(sp.n), kFxBlock, 0, s>>>(st, sp, bound, cell_bound);
  k_fx_p2g_settle<<<1, 256, 0, s>>>(st, sp, grid, bound, cell_bound, ncells, counters);
  return hipGetLastError();
}

hipError_t fx_launch_grid(int64_t* grid, uint32_t ncells, const int64_t gdt[3], unsigned long long* bound,
                          hipStream_t s) {
  k_fx_grid<<<blocks_for(ncells), kFxBlock, 0, s>>>(grid, ncells, gdt[0], gdt[1], gdt[2], bound);
  return hipGetLastError();
}

hipError_t fx_launch_g2p(int model, const FxStore& st, const FxSubstep& sp, const int64_t* grid,
                         const int64_t* sin_tab, const FxSortBufs* sort, hipStream_t s) {
  if (!sp.n) return hipSuccess;
  const unsigned g = blocks_for(sp.n);
  FxOrder ord{nullptr, nullptr, nullptr, 0};
  if (sort) ord = FxOrder{sort->rank, sort->rec[sort->cur], sort->keys[sort->cur], sort->nbins - 1};
  switch (model) {
    case FXM_ELASTIC: k_fx_g2p<FXM_ELASTIC><<<g, kFxBlock, 0, s>>>(st, sp, grid, sin_tab, ord); break;
    case FXM_ELASTOPLASTIC: k_fx_g2p<FXM_ELASTOPLASTIC><<<g, kFxBlock, 0, s>>>(st, sp, grid, sin_tab, ord); break;
    case FXM_GRANULAR: k_fx_g2p<FXM_GRANULAR><<<g, kFxBlock, 0, s>>>(st, sp, grid, sin_tab, ord); break;
    case FXM_BRITTLE: k_fx_g2p<FXM_BRITTLE><<<g, kFxBlock, 0, s>>>(st, sp, grid, sin_tab, ord); break;
    case FXM_FLUID: k_fx_g2p<FXM_FLUID><<<g, kFxBlock, 0, s>>>(st, sp, grid, sin_tab, ord); break;
    case FXM_SNOW: k_fx_g2p<FXM_SNOW><<<g, kFxBlock, 0, s>>>(st, sp, grid, sin_tab, ord); break;
    default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t fx_launch_frac_mark(const FxStore& st, uint32_t n, int64_t threshold, uint64_t* key, int64_t* strain,
                               hipStream_t s) {
  if (!n) return hipSuccess;
  k_fx_frac_mark<<<blocks_for(n), kFxBlock, 0, s>>>(st, n, threshold, key, strain);
  return hipGetLastError();
}

size_t fx_scan_temp_bytes(uint32_t n) {
  size_t bytes = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr, (int)n);
  return bytes;
}

hipError_t fx_launch_scan(const uint64_t* key, uint64_t* off, uint32_t n, void* temp, size_t temp_bytes,
                          uint64_t* total_host, hipStream_t s) {
  if (n) {
    hipError_t e = hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, key, off, (int)n, s);
    if (e != hipSuccess) return e;
  }
  k_fx_total<<<1, 64, 0, s>>>(key, off, n, total_host);
  return hipGetLastError();
}

hipError_t fx_launch_frac_spawn(const FxStore& st, uint32_t n, const uint64_t* key, const uint64_t* off,
                                const int64_t* strain, const int64_t* frag_off, uint64_t next_id, double time,
                                FxEvent* events, hipStream_t s) {
  if (!n) return hipSuccess;
  k_fx_frac_spawn<<<blocks_for(n), kFxBlock, 0, s>>>(st, n, key, off, strain, frag_off, next_id, time, events);
  return hipGetLastError();
}

hipError_t fx_launch_put(const FxStore& st, uint32_t at, const uint64_t* src, uint32_t m, hipStream_t s) {
  if (!m) return hipSuccess;
  k_fx_put<<<blocks_for((size_t)m * FX_NPLANES), kFxBlock, 0, s>>>(st, at, src, m);
  return hipGetLastError();
}

hipError_t fx_launch_get(const FxStore& st, uint32_t first, uint32_t m, uint64_t* dst, hipStream_t s) {
  if (!m) return hipSuccess;
  k_fx_get<<<blocks_for((size_t)m * FX_NPLANES), kFxBlock, 0, s>>>(st, first, m, dst);
  return hipGetLastError();
}

hipError_t fx_launch_find(const FxStore& st, uint32_t n, uint64_t id, uint32_t* idx, hipStream_t s) {
  if (!n) return hipSuccess;
  k_fx_find<<<blocks_for(n), kFxBlock, 0, s>>>(st, n, id, idx);
  return hipGetLastError();
}

hipError_t fx_launch_stats(const FxStore& st, uint32_t n, FxStatsAcc* acc, hipStream_t s) {
  if (!n) return hipSuccess;
  k_fx_stats<<<blocks_for(n), kFxBlock, 0, s>>>(st, n, acc);
  return hipGetLastError();
}

hipError_t fx_launch_stats_serial(const FxStore& st, uint32_t n, FxStatsAcc* acc, hipStream_t s) {
  k_fx_stats_serial<<<1, 64, 0, s>>>(st, n, acc);
  return hipGetLastError();
}

hipError_t fx_launch_snapshots(const FxStore& st, uint32_t n, FxSnapshot* out, hipStream_t s) {
  if (!n) return hipSuccess;
  k_fx_snap<<<blocks_for(n), kFxBlock, 0, s>>>(st, n, out);
  return hipGetLastError();
}
