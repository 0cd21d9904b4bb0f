// Entry kernels and launchers of the FP64 step kernel (one box with any
// collision rule and the rigid sphere, or every box of an SRD ensemble in
// one launch); the tile code is in mpcd_tile.hpp.
#include "mpcd_tile.hpp"

namespace mcm {

template <int TX, int TY, int TZ, int K, bool TRACK, bool BEADS, int RULE, bool SPHERE>
__global__ __launch_bounds__(kBlock) void k_step_f64(StepArgs A)
{
    step_f64_tile<TX, TY, TZ, K, TRACK, BEADS, RULE, SPHERE, false>(A, blockIdx.x);
}

template <int TX, int TY, int TZ, int K, bool TRACK, bool BEADS>
__global__ __launch_bounds__(kBlock) void k_step_f64_multi(const StepArgs *__restrict__ args,
                                                          const int *__restrict__ tile_start, int nbox)
{
    const int b = box_of_block(tile_start, nbox);
    step_f64_tile<TX, TY, TZ, K, TRACK, BEADS, kRuleSrd, false, false>(args[b], (int)blockIdx.x - tile_start[b]);
}

template <int V, int RULE, bool SPHERE>
static void launch_f64_v(const StepArgs &A, bool track, bool beads, hipStream_t s)
{
    constexpr TileVariant t = kF64Variants[V];
    dispatch_tb(track, beads, [&](auto tr, auto bd) {
        hipLaunchKernelGGL((k_step_f64<t.tx, t.ty, t.tz, t.k, decltype(tr)::value, decltype(bd)::value, RULE, SPHERE>),
                           dim3(A.g.ntiles), dim3(kBlock), 0, s, A);
    });
}

template <int V>
static void launch_f64_multi_v(const StepArgs *args, const int *tile_start, int nbox, int ntiles, bool track,
                               bool beads, hipStream_t s)
{
    constexpr TileVariant t = kF64Variants[V];
    dispatch_tb(track, beads, [&](auto tr, auto bd) {
        hipLaunchKernelGGL((k_step_f64_multi<t.tx, t.ty, t.tz, t.k, decltype(tr)::value, decltype(bd)::value>),
                           dim3(ntiles), dim3(kBlock), 0, s, args, tile_start, nbox);
    });
}

void launch_step_f64(const StepArgs &A, int variant, bool track, bool beads, bool deferred, hipStream_t s)
{
    static_assert(kNumF64Variants == 2, "launcher table");
    if (deferred) {
        launch_deferred_f64(A, variant, track, beads, s);
        return;
    }
    // MPC-AT keeps two more per-particle arrays in LDS: the small tile only
    // the rigid sphere (SRD only) has its own instantiations, so runs
    // without it keep the lean code
    if (A.rule == kRuleAt)
        launch_f64_v<kAtVariant, kRuleAt, false>(A, track, beads, s);
    else if (A.rule == kRuleAtA)
        launch_f64_v<kAtVariant, kRuleAtA, false>(A, track, beads, s);
    else if (variant == 1)
        A.sphere ? launch_f64_v<1, kRuleSrd, true>(A, track, beads, s) : launch_f64_v<1, kRuleSrd, false>(A, track, beads, s);
    else
        A.sphere ? launch_f64_v<0, kRuleSrd, true>(A, track, beads, s) : launch_f64_v<0, kRuleSrd, false>(A, track, beads, s);
}

void launch_step_multi_f64(const StepArgs *args, const int *tile_start, int nbox, int ntiles, int variant, bool track,
                           bool beads, hipStream_t s)
{
    if (variant == 1)
        launch_f64_multi_v<1>(args, tile_start, nbox, ntiles, track, beads, s);
    else
        launch_f64_multi_v<0>(args, tile_start, nbox, ntiles, track, beads, s);
}

}  // namespace mcm
