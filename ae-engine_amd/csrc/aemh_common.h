// aemh_common.h -- POD structures shared by host code and kernels.
#pragma once
#include <stdint.h>

namespace aemh {

constexpr int kModelNeoHookean = 0;
constexpr int kModelDruckerPrager = 1;
constexpr int kModelVonMises = 2;
constexpr int kModelNewtonianFluid = 3;
constexpr int kModelCorotated = 4;
constexpr int kModelFixedCorotated = 5;
constexpr int kModelRigid = 6;
constexpr int kModelSnow = 7;

// Grid is stored block-major: 4x4x4 nodes per block, 64 consecutive nodes
// (x fastest inside a block), blocks ordered z-y-x over each scene's padded
// grid and scenes concatenated.  Only blocks touched by particles in the
// current substep are cleared, updated and read.
constexpr int kBlockNodes = 64;
// P2G stages the 6x6x6 node halo of a 4x4x4 block of base cells in LDS.
constexpr int kTileSide = 6;
constexpr int kTileNodes = 216;
constexpr int kMaxTileBudget = 5;  // most LDS tiles a P2G workgroup may use

// One scene as seen by the kernels (derived from aemh_scene + aemh_config).
struct SceneDev {
  uint32_t first, count;  // particle range in sorted (working) order
  int32_t n[3];           // grid nodes
  int32_t nb[3];          // blocks per axis = ceil(n/4)
  uint32_t block_off;     // first global block of this scene
  float dx, inv_dx;
  float o[3];             // moving-window origin
  float dt;               // substep dt
  float g[3];             // gravity
  float damping;
  int32_t bc_margin;
  float bc_rest;
  uint32_t vel_only;      // velocity-only tier semantics
  uint32_t window_clamp;  // 1: origin-relative x/z clamp
  float gmax_x, gmax_z;   // (n-1) dx
  float four_inv_dx2;     // 4 inv_dx inv_dx, evaluated left to right
  uint32_t pad;
};

// Material as the kernels need it.  Derived constants are computed on the
// host with the reference's f32 operation order (constitutive.rs:55-82).
struct MatDev {
  float mu, lambda, bulk, density;
  float yield_stress, dp_alpha, dp_k, pad0;
  float theta_c, theta_s, xi, pad1;
  int32_t model;
  int32_t pad2[3];
};

// Particle state in HBM, structure of arrays.  C and F may be stored in
// 16-bit floats (F as F - I) in the mixed-precision mode.
struct PBuf {
  float* x[3];
  float* v[3];
  float* C[9];
  float* F[9];
  // 16-bit storage (mixed precision): two 16-bit values per 32-bit word.
  // Words 0-3 hold C pairs (C[2w] in the low half, C[2w+1] in the high),
  // words 4-7 the pairs of F - I, word 8 C[8] (low) and F[8] - 1 (high).
  uint32_t* CF[9];
  float* J;
  float* Jp;
  float* mass;
  uint16_t* mat;
  uint32_t* id;     // position in the active index list
};

// Lifecycle fields of a resident MpssBuffer (mpss.rs:104-157).
struct MpssDev {
  float* lifetime;
  float* age;
  uint8_t* active;
};

// Up to 9 component arrays of one SoA field.
struct Ptr9 {
  float* p[9];
};

// Grid accumulators and the updated velocity read by G2P.
struct GridDev {
  float* accf;      // [7][nodes] m, p0..2, f0..2 (FP32 mode)
  long long* accq;  // [8][nodes] Q32.32 raw sums (deterministic mode); field 7
                    // bounds sum |contribution| in units of 2^32 raw
  float4* gv;       // [nodes] (v.xyz, m) after the grid update
  // P2G tile partials, [workgroup][tile][field][216 tile nodes]: FP32 sums
  // (FP32 mode) or Q32.32 raw sums (deterministic mode)
  float* partf;
  long long* partq;
  uint32_t nodes;
};
constexpr int kDetFields = 8;
// Accumulator planes of the grid, by mode (numbered as AEMH_ACCUM_*): the
// FP32 mode's 4 f32 planes (m, p + f dt), the deterministic mode's 8 int64
// planes, the reference-order mode's 7 f32 planes (m, p, f).
constexpr int kAccF32Fold = 0, kAccQ32 = 1, kAccF32Mpf = 2;
// counters[0]: active blocks of the substep; counters[4]: sticky Q32.32
// overflow flag (a node's sum |contribution| reached 2^63 raw).
constexpr int kCntActive = 0;
constexpr int kCntOverflow = 4;
// counters[5]: snow SVDs that took the FP64 path since the particles were
// uploaded (cumulative; the scan clears only counters[0..3])
constexpr int kCntSvd64 = 5;

// Per-substep bookkeeping buffers.
struct SortDev {
  uint32_t* key;        // [N] global base-block of each working slot
  uint32_t* rank;       // [N] rank within its block
  uint32_t* perm;       // [N] sorted position -> working slot
  uint32_t* cnt;        // [NB] particles per base block (zeroed by the scan)
  uint32_t* off;        // [NB] tile-local exclusive prefix
  uint32_t* traw;       // [tiles] tile sums of the count scan
  uint32_t* tsum;       // [tiles] their exclusive scan
  uint32_t* flag;       // [NB] node block active this substep
  uint32_t* active;     // [NB] list of active node blocks
  uint32_t* counters;   // [8], see kCnt*
  uint32_t* runs;       // [64 x 16] wave-run counts of the count kernel, one
                        // per 64 B line (kRunLanes counters, stride kRunStride)
  uint32_t* runs_host;  // device address of a host-mapped word: the runs of
                        // the last count (particles' block-order disorder)
  uint32_t* bslot;      // [NB] tile slot of the block in the P2G workgroup
                        // holding its first particle (>= tiles: no tile)
};

// The float range of the host logf table of the reference-order FP32 mode:
// [0.1, 10], where G2P clamps J (mpm_solver.rs:1040-1042).  Positive floats
// order as their bit patterns, so entry i is logf of the float with bits
// kLnTabLo + i.
constexpr uint32_t kLnTabLo = 0x3DCCCCCDu;  // bits of 0.1f
constexpr uint32_t kLnTabHi = 0x41200000u;  // bits of 10.0f
constexpr uint32_t kLnTabSize = kLnTabHi - kLnTabLo + 1u;

// Reference-order FP32 P2G (AEMH_ACCUM_F32_SERIAL): every visited stencil
// node of every working slot is an entry, sorted by (node, list position)
// so each node's sums run in the reference's particle order.
struct SerialDev {
  unsigned long long* key[2];  // [27 N] (node << id_bits) | list position; unused: nodes << id_bits
  uint32_t* val[2];            // [27 N] entry slot 27 j + k
  float* contrib;              // [27 N][8] m, p0..2, f0..2 (pad) of each entry
  const float* lntab;          // host logf of every float in [0.1, 10]
  void* tmp;                   // radix-sort temporary storage
  size_t tmp_bytes;
  int id_bits, end_bit;        // bits of the list position; key bits to sort
};

constexpr int kScanTile = 1024;
constexpr int kRunLanes = 64;
constexpr int kRunStride = 16;

}  // namespace aemh
