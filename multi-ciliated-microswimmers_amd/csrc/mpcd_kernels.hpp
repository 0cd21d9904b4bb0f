// Tile-fused MPCD step kernels for gfx950.
//
// The box is cut into tiles of TX x TY x TZ collision cells (in the frame of
// the current random grid shift).  Each tile's particles sit in a "home"
// segment plus the 26 "outbox" segments its neighbours wrote for it during
// the previous step.  One workgroup per tile:
//   load (home + 26 inbound outboxes) -> per-cell sums in LDS -> per-cell
//   thermostat + rotation -> per-particle update -> stream -> new cell under
//   the next shift -> write to the new home or to an outbox.
// Every particle record is read once and written once per step.
#pragma once

#include "mpcd_device.hpp"

namespace mcm {

constexpr int kBlock = 256;

// Tile shapes (cells) and per-thread particle slots K of the step kernel
// variants (256 threads; capacity K * 256 particles per tile).  The host
// picks one per context (fixed, or timed by the autotuner).
struct TileVariant {
    int tx, ty, tz, k;
};
constexpr int kNumF32Variants = 3;
constexpr TileVariant kF32Variants[kNumF32Variants] = {{8, 4, 4, 7}, {4, 4, 4, 3}, {4, 4, 8, 7}};
constexpr int kNumF64Variants = 2;
constexpr TileVariant kF64Variants[kNumF64Variants] = {{8, 4, 4, 6}, {4, 4, 4, 3}};

// Tiles above a variant's capacity (k particles per thread) are stepped by
// a second pass holding deferred_k per thread: 32 particles per cell of
// the tile.  Home segments hold that many too.
constexpr int kDeferredPerCell = 32;
constexpr int deferred_k(const TileVariant &t) { return (t.tx * t.ty * t.tz * kDeferredPerCell + 255) / 256; }

struct Geom {
    int L[3];
    double invL[3];  // 1 / L (box_axis)
    int T[3];
    int nt[3];
    int ntiles;
    int cap_home;
    int ob_stride;
    int ob_off[26];
    int ob_cap[26];
    int cap_ovf;
};

// Per-step diagnostics are accumulated by each workgroup into one of
// kDiagSlots slots (tile index modulo kDiagSlots), so no single address sees
// every workgroup's atomics; k_begin_step folds the slots into the running
// totals and clears them.
constexpr int kDiagSlots = 256;

struct DiagSlot {
    long long q_mom[3];  // FP32-bulk: exact momentum in 2^-21 units
    long long q_diss;    // FP32-bulk: fluid dissipated energy in 2^-42 units
    double mom[3];
    double e2;           // sum m v^2 after
    double diss;         // FP64 part of the dissipated energy
    long long n;
    long long rel_viol;
    int max_load;        // most particles one tile of this slot held
    int pad;
};

// The step runs in up to three passes over a box's tiles, all with the
// same tile code: the main pass over every tile (without the bead code),
// the bead pass over the tiles the main pass found holding beads, and the
// large pass, with a larger capacity, over tiles above the capacity of the
// other two (at most kDeferCap per step).
enum : int { kPassMain = 0, kPassBeads = 1, kPassLarge = 2 };
constexpr int kDeferCap = 4096;
constexpr int kDeferBlocks = 64;
constexpr int kBeadPassBlocks = 1024;   // one box; an ensemble: kEnsBeadBlocks per box
constexpr int kEnsBeadBlocks = 64;

struct Diag {
    DiagSlot slot[kDiagSlots];
    long long run_q_diss;
    double run_diss;
    long long run_rel_viol;
    long long pad;
    int defer_n;         // tiles of this step deferred to the large pass
    int bead_n;          // tiles of this step holding beads (StepArgs::bead_tiles)
    int defer_tile[kDeferCap];
};

enum : int {
    kFlagTileCap = 1,      // a tile held more particles than the kernel's capacity
    kFlagOverflowCap = 2,  // overflow list full
    kFlagHomeCap = 4,      // overflow fix-up found a full home segment
    kFlagVelocity = 8,     // FP32-bulk velocity beyond the exact range
    kFlagMisbinned = 16,   // FP64 particle found outside its tile
    kFlagBeadTile = 32,    // bead outside the tile range
};

enum : int { kRuleSrd = 0, kRuleAt = 1, kRuleAtA = 2 };

// Solid objects the FP32-bulk tile code takes (an instantiation each): none,
// the rigid sphere, the planar walls.
enum : int { kObjNone = 0, kObjSphere = 1, kObjWalls = 2 };

// A rigid sphere in device memory: state, and the impulse of the running
// step (virtual particles in the collision, bounce-back in the stream),
// applied by k_sphere_finish after the step.
struct SphereDev {
    double R, Ms, Is;
    double X[3], V[3], W[3];
    int slip, pad;
    double dP[3], dL[3];        // impulse of the running step
    double lastP[3], lastL[3];  // impulse of the last completed step
    double h_last;              // its time step
};
constexpr int kMaxSphereCand = 64;  // virtual-particle candidates per cell (rho)
// tile variant of the MPC-AT rules in both forms (4 x 4 x 4, three
// particles per thread: its LDS arrays per particle fit four tiles per CU
// in the FP32-bulk form)
constexpr int kAtVariant = 1;
// FP64 SRD default: the 4 x 4 x 4 tile holding 3 particles per thread, 1.2
// times the mean load at rho 10 (five tiles per CU fit its LDS and
// registers: 5.20 ms at 256^3, against 5.42 ms holding 4 per thread and
// 6.0 ms for the 8 x 4 x 4 tile)
constexpr int kF64DefaultVariant = 1;

struct StepArgs {
    Geom g;
    const uint32_t *in_home;
    const uint32_t *in_ob;
    const int *in_home_cnt;
    const int *in_ob_cnt;
    uint32_t *out_home;
    uint32_t *out_ob;
    int *out_home_cnt;
    int *out_ob_cnt;
    uint32_t *ovf;
    int *ovf_cnt;       // counter of this step's overflow list
    int *flags;
    double a[3], a_next[3];
    float da[3];
    double h;
    float hf;
    double calpha, salpha, kT;
    int thermostat;
    uint64_t seed, step;
    const double *inj_gamma;
    const double *inj_axis;
    // planar walls at z = 0 and z = L[2] - 1
    int walls;           // kWallsNone, kWallsNoSlip, kWallsSlip
    int rho;             // wall-layer cells below rho get rho - num virtual particles
    double vw[2];        // top wall velocity (NOSLIP)
    const double *inj_virt;  // 3 normals per cell replacing the Philox draws
    // collision rule (kRule*); MPC-AT: injected normals, 3 per particle id
    // (beads after the nf_total solvent ids), bead positions (staged order)
    int rule;
    const double *inj_ran;
    long long nf_total;
    const double *b_pos3;
    // rigid sphere; injected normals, 3 per (cell, candidate)
    SphereDev *sphere;
    const double *inj_sph;
    // beads, sorted by (tile, local cell, index)
    int *bead_tiles;     // the bead pass's tiles (capacity: the box's tiles)
    int nb;
    const int *b_tile_start;
    const int *b_lc;
    const int *b_idx;
    const double *b_vel4;
    double *b_out4;
    Diag *diag;
    int *dbg_cell;
    int *dbg_rank;
};

// ---------------------------------------------------------------- records
// Records live in device (global) memory; their base pointers reach the
// tile code through LDS tables, which hides that from the compiler, so the
// accesses go through global-address-space pointers: global loads and
// stores (merged into dwordx4 + dwordx3), not flat ones, whose completion
// every later wait on an LDS operation would otherwise include.
typedef __attribute__((address_space(1))) uint32_t global_u32;

template <int W>
__device__ __forceinline__ void ld_words(const uint32_t *__restrict__ p, uint32_t (&o)[W])
{
    const global_u32 *g = (const global_u32 *)p;
#pragma unroll
    for (int i = 0; i < W; ++i)
        o[i] = g[i];
}

template <int W>
__device__ __forceinline__ void st_words(uint32_t *__restrict__ p, const uint32_t (&o)[W])
{
    global_u32 *g = (global_u32 *)p;
#pragma unroll
    for (int i = 0; i < W; ++i)
        g[i] = o[i];
}

__device__ __forceinline__ uint32_t fbits(float f) { return __float_as_uint(f); }
__device__ __forceinline__ float bitsf(uint32_t u) { return __uint_as_float(u); }

__device__ __forceinline__ void dsplit(double d, uint32_t &lo, uint32_t &hi)
{
    const uint64_t u = (uint64_t)__double_as_longlong(d);
    lo = (uint32_t)u;
    hi = (uint32_t)(u >> 32);
}
__device__ __forceinline__ double djoin(uint32_t lo, uint32_t hi)
{
    return __longlong_as_double((long long)(((uint64_t)hi << 32) | lo));
}

// packed cell key of the FP32-bulk record: 10 bits per axis
__host__ __device__ __forceinline__ uint32_t pack_cell(int cx, int cy, int cz)
{
    return (uint32_t)cx | ((uint32_t)cy << 10) | ((uint32_t)cz << 20);
}

// ------------------------------------------------------- segment table
struct TileCoord {
    int t, tx, ty, tz, ox, oy, oz;
};

__device__ __forceinline__ TileCoord tile_coord(const Geom &g, int t)
{
    TileCoord c;
    c.t = t;
    c.tx = t % g.nt[0];
    c.ty = (t / g.nt[0]) % g.nt[1];
    c.tz = t / (g.nt[0] * g.nt[1]);
    c.ox = c.tx * g.T[0];
    c.oy = c.ty * g.T[1];
    c.oz = c.tz * g.T[2];
    return c;
}

__device__ __forceinline__ int wrap_t(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// Segment 0 = home of the tile; segment 1 + d = outbox d of the neighbour
// the direction d points away from.  Writes count and record base.
__device__ __forceinline__ void segment_info(const Geom &g, const TileCoord &c, int s, const int *home_cnt,
                                             const int *ob_cnt, int &cnt, long long &base)
{
    if (s == 0) {
        cnt = min(home_cnt[c.t], g.cap_home);
        base = (long long)c.t * g.cap_home;
        return;
    }
    const int d = s - 1;
    const int i = d < 13 ? d : d + 1;
    const int dx = i / 9 - 1, dy = (i / 3) % 3 - 1, dz = i % 3 - 1;
    const int sx = wrap_t(c.tx - dx, g.nt[0]);
    const int sy = wrap_t(c.ty - dy, g.nt[1]);
    const int sz = wrap_t(c.tz - dz, g.nt[2]);
    const int src = sx + g.nt[0] * (sy + g.nt[1] * sz);
    cnt = min(ob_cnt[src * 26 + d], g.ob_cap[d]);
    base = (long long)src * g.ob_stride + g.ob_off[d];
}

// Wave-aggregated LDS counter bump for lanes with pred set.
__device__ __forceinline__ int wave_bump(int *ctr, bool pred)
{
    const uint64_t m = __ballot(pred);
    if (m == 0)
        return 0;
    const int leader = __ffsll((long long)m) - 1;
    int base = 0;
    if (lane_id() == leader)
        base = atomicAdd(ctr, __popcll(m));
    base = __builtin_amdgcn_readlane(base, leader);
    return base + (int)mbcnt64(m);
}

// One pass (kPass*) of the step over one box; the passes run in order, the
// bead pass only when the box has beads, then the overflow fix-up.
void launch_step_f32(const StepArgs &A, int variant, bool track, int pass, hipStream_t s);
void launch_step_f64(const StepArgs &A, int variant, bool track, int pass, hipStream_t s);
// One pass over every box of an ensemble (device arrays: per-box arguments
// and, for the main pass, the first workgroup of each box, tile_start[nbox]
// = ntiles).
// rule: the boxes' collision rule (MPC-AT boxes use kAtVariant); sphere:
// every box has a rigid sphere (SRD; FP32-bulk: tile variant 0).
void launch_step_multi(bool f32, const StepArgs *args, const int *tile_start, int nbox, int ntiles, int variant,
                       bool track, int pass, hipStream_t s, int rule = kRuleSrd, bool sphere = false);
// the pieces of the above, one per translation unit
void launch_step_multi_f64(const StepArgs *args, const int *tile_start, int nbox, int ntiles, int variant, bool track,
                           hipStream_t s, int rule = kRuleSrd, bool sphere = false);
void launch_later_f32(const StepArgs &A, int variant, bool track, int pass, hipStream_t s);
void launch_later_f64(const StepArgs &A, int variant, bool track, int pass, hipStream_t s);
void launch_later_multi(bool f32, const StepArgs *args, int nbox, int variant, bool track, int pass, hipStream_t s,
                        int rule = kRuleSrd, bool sphere = false);
void launch_later_multi_f64(const StepArgs *args, int nbox, int variant, bool track, int pass, hipStream_t s,
                            int rule = kRuleSrd, bool sphere = false);

}  // namespace mcm
