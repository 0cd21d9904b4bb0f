// Entry kernels and launchers of the FP32-bulk step kernel (one box, or
// every box of an ensemble in one launch); the tile code is in mpcd_tile.hpp.
#include "mpcd_tile.hpp"

namespace mcm {

template <int TX, int TY, int TZ, int K, bool TRACK, int RULE, int OBJ = kObjNone>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_step_f32(StepArgs A)
{
    step_f32_tile<TX, TY, TZ, K, TRACK, false, kPassMain, RULE, OBJ>(A, blockIdx.x);
}

template <int TX, int TY, int TZ, int K, bool TRACK, int RULE = kRuleSrd, int OBJ = kObjNone>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) void k_step_f32_multi(
    const StepArgs *__restrict__ args, const int *__restrict__ tile_start, int nbox)
{
    const int b = box_of_block(tile_start, nbox);
    step_f32_tile<TX, TY, TZ, K, TRACK, false, kPassMain, RULE, OBJ>(args[b], (int)blockIdx.x - tile_start[b]);
}

template <int V, int RULE = kRuleSrd, int OBJ = kObjNone>
static void launch_f32_v(const StepArgs &A, bool track, hipStream_t s)
{
    constexpr TileVariant t = kF32Variants[V];
    if (track)
        hipLaunchKernelGGL((k_step_f32<t.tx, t.ty, t.tz, t.k, true, RULE, OBJ>), dim3(A.g.ntiles), dim3(kBlock), 0, s,
                           A);
    else
        hipLaunchKernelGGL((k_step_f32<t.tx, t.ty, t.tz, t.k, false, RULE, OBJ>), dim3(A.g.ntiles), dim3(kBlock), 0, s,
                           A);
}

template <int V, int RULE = kRuleSrd, int OBJ = kObjNone>
static void launch_f32_multi_v(const StepArgs *args, const int *tile_start, int nbox, int ntiles, bool track,
                               hipStream_t s)
{
    constexpr TileVariant t = kF32Variants[V];
    if (track)
        hipLaunchKernelGGL((k_step_f32_multi<t.tx, t.ty, t.tz, t.k, true, RULE, OBJ>), dim3(ntiles), dim3(kBlock), 0,
                           s, args, tile_start, nbox);
    else
        hipLaunchKernelGGL((k_step_f32_multi<t.tx, t.ty, t.tz, t.k, false, RULE, OBJ>), dim3(ntiles), dim3(kBlock), 0,
                           s, args, tile_start, nbox);
}

void launch_step_f32(const StepArgs &A, int variant, bool track, int pass, hipStream_t s)
{
    static_assert(kNumF32Variants == 3, "launcher table");
    if (pass != kPassMain) {
        launch_later_f32(A, variant, track, pass, s);
        return;
    }
    // a rigid sphere or planar walls: their instantiations of the default
    // tile (the context uses variant 0)
    if (A.sphere) {
        launch_f32_v<0, kRuleSrd, kObjSphere>(A, track, s);
        return;
    }
    if (A.walls) {
        launch_f32_v<0, kRuleSrd, kObjWalls>(A, track, s);
        return;
    }
    // MPC-AT keeps a random velocity per slot and 15 integer sums per cell
    // in LDS: the small tile
    if (A.rule == kRuleAt) {
        launch_f32_v<kAtVariant, kRuleAt>(A, track, s);
        return;
    }
    if (A.rule == kRuleAtA) {
        launch_f32_v<kAtVariant, kRuleAtA>(A, track, s);
        return;
    }
    switch (variant) {
    case 1: launch_f32_v<1>(A, track, s); break;
    case 2: launch_f32_v<2>(A, track, s); break;
    default: launch_f32_v<0>(A, track, s); break;
    }
}

void launch_step_multi(bool f32, const StepArgs *args, const int *tile_start, int nbox, int ntiles, int variant,
                       bool track, int pass, hipStream_t s, int rule, bool sphere)
{
    if (pass != kPassMain) {
        launch_later_multi(f32, args, nbox, variant, track, pass, s, rule, sphere);
        return;
    }
    if (!f32) {
        launch_step_multi_f64(args, tile_start, nbox, ntiles, variant, track, s, rule, sphere);
        return;
    }
    if (sphere) {  // the rigid sphere's instantiations of the default tile
        launch_f32_multi_v<0, kRuleSrd, kObjSphere>(args, tile_start, nbox, ntiles, track, s);
        return;
    }
    if (rule == kRuleAt) {
        launch_f32_multi_v<kAtVariant, kRuleAt>(args, tile_start, nbox, ntiles, track, s);
        return;
    }
    if (rule == kRuleAtA) {
        launch_f32_multi_v<kAtVariant, kRuleAtA>(args, tile_start, nbox, ntiles, track, s);
        return;
    }
    switch (variant) {
    case 1: launch_f32_multi_v<1>(args, tile_start, nbox, ntiles, track, s); break;
    case 2: launch_f32_multi_v<2>(args, tile_start, nbox, ntiles, track, s); break;
    default: launch_f32_multi_v<0>(args, tile_start, nbox, ntiles, track, s); break;
    }
}

}  // namespace mcm
