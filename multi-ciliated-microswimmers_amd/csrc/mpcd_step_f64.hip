// Entry kernels and launchers of the FP64 step kernel (one box with any
// collision rule and the rigid sphere, or every box of an SRD ensemble in
// one launch); the tile code is in mpcd_tile.hpp.
#include "mpcd_tile.hpp"

namespace mcm {

template <int TX, int TY, int TZ, int K, bool TRACK, int RULE, bool SPHERE>
__global__ __launch_bounds__(kBlock) void k_step_f64(StepArgs A)
{
    step_f64_tile<TX, TY, TZ, K, TRACK, false, RULE, SPHERE, kPassMain>(A, blockIdx.x);
}

template <int TX, int TY, int TZ, int K, bool TRACK, int RULE = kRuleSrd, bool SPHERE = false>
__global__ __launch_bounds__(kBlock) void k_step_f64_multi(const StepArgs *__restrict__ args,
                                                          const int *__restrict__ tile_start, int nbox)
{
    const int b = box_of_block(tile_start, nbox);
    step_f64_tile<TX, TY, TZ, K, TRACK, false, RULE, SPHERE, kPassMain>(args[b], (int)blockIdx.x - tile_start[b]);
}

template <int V, int RULE, bool SPHERE>
static void launch_f64_v(const StepArgs &A, bool track, hipStream_t s)
{
    constexpr TileVariant t = kF64Variants[V];
    if (track)
        hipLaunchKernelGGL((k_step_f64<t.tx, t.ty, t.tz, t.k, true, RULE, SPHERE>), dim3(A.g.ntiles), dim3(kBlock), 0,
                           s, A);
    else
        hipLaunchKernelGGL((k_step_f64<t.tx, t.ty, t.tz, t.k, false, RULE, SPHERE>), dim3(A.g.ntiles), dim3(kBlock),
                           0, s, A);
}

template <int V, int RULE = kRuleSrd, bool SPHERE = false>
static void launch_f64_multi_v(const StepArgs *args, const int *tile_start, int nbox, int ntiles, bool track,
                               hipStream_t s)
{
    constexpr TileVariant t = kF64Variants[V];
    if (track)
        hipLaunchKernelGGL((k_step_f64_multi<t.tx, t.ty, t.tz, t.k, true, RULE, SPHERE>), dim3(ntiles), dim3(kBlock),
                           0, s, args, tile_start, nbox);
    else
        hipLaunchKernelGGL((k_step_f64_multi<t.tx, t.ty, t.tz, t.k, false, RULE, SPHERE>), dim3(ntiles), dim3(kBlock),
                           0, s, args, tile_start, nbox);
}

void launch_step_f64(const StepArgs &A, int variant, bool track, int pass, hipStream_t s)
{
    static_assert(kNumF64Variants == 2, "launcher table");
    if (pass != kPassMain) {
        launch_later_f64(A, variant, track, pass, s);
        return;
    }
    // MPC-AT keeps two more per-particle arrays in LDS: the small tile only
    // the rigid sphere (SRD only) has its own instantiations, so runs
    // without it keep the lean code
    if (A.rule == kRuleAt)
        launch_f64_v<kAtVariant, kRuleAt, false>(A, track, s);
    else if (A.rule == kRuleAtA)
        launch_f64_v<kAtVariant, kRuleAtA, false>(A, track, s);
    else if (variant == 1)
        A.sphere ? launch_f64_v<1, kRuleSrd, true>(A, track, s) : launch_f64_v<1, kRuleSrd, false>(A, track, s);
    else
        A.sphere ? launch_f64_v<0, kRuleSrd, true>(A, track, s) : launch_f64_v<0, kRuleSrd, false>(A, track, s);
}

void launch_step_multi_f64(const StepArgs *args, const int *tile_start, int nbox, int ntiles, int variant, bool track,
                           hipStream_t s, int rule, bool sphere)
{
    if (sphere && variant == 1)
        launch_f64_multi_v<1, kRuleSrd, true>(args, tile_start, nbox, ntiles, track, s);
    else if (sphere)
        launch_f64_multi_v<0, kRuleSrd, true>(args, tile_start, nbox, ntiles, track, s);
    else if (rule == kRuleAt)
        launch_f64_multi_v<kAtVariant, kRuleAt>(args, tile_start, nbox, ntiles, track, s);
    else if (rule == kRuleAtA)
        launch_f64_multi_v<kAtVariant, kRuleAtA>(args, tile_start, nbox, ntiles, track, s);
    else if (variant == 1)
        launch_f64_multi_v<1>(args, tile_start, nbox, ntiles, track, s);
    else
        launch_f64_multi_v<0>(args, tile_start, nbox, ntiles, track, s);
}

}  // namespace mcm
