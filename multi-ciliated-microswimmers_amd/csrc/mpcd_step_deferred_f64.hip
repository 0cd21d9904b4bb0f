// The deferred pass of the FP64 step: tiles above the step kernel's
// capacity, stepped by the same tile code with the larger capacity
// deferred_k (kDeferBlocks workgroups per box, striding over the box's
// deferred list).
#include "mpcd_tile.hpp"

namespace mcm {

template <int TX, int TY, int TZ, int KB, bool TRACK, bool BEADS, int RULE, bool SPHERE>
__device__ __forceinline__ void f64_deferred_tiles(const StepArgs &A)
{
    const int n = min(A.diag->defer_n, kDeferCap);
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        step_f64_tile<TX, TY, TZ, KB, TRACK, BEADS, RULE, SPHERE, true>(A, A.diag->defer_tile[i]);
        __syncthreads();
    }
}

template <int TX, int TY, int TZ, int KB, bool TRACK, bool BEADS, int RULE, bool SPHERE>
__global__ __launch_bounds__(kBlock) void k_step_f64_deferred(StepArgs A)
{
    f64_deferred_tiles<TX, TY, TZ, KB, TRACK, BEADS, RULE, SPHERE>(A);
}

// blockIdx.y = box (ensembles: SRD, no sphere)
template <int TX, int TY, int TZ, int KB, bool TRACK, bool BEADS>
__global__ __launch_bounds__(kBlock) void k_step_f64_deferred_multi(const StepArgs *__restrict__ args)
{
    f64_deferred_tiles<TX, TY, TZ, KB, TRACK, BEADS, kRuleSrd, false>(args[blockIdx.y]);
}

template <int V, int RULE, bool SPHERE>
static void launch_v(const StepArgs &A, bool track, bool beads, hipStream_t s)
{
    constexpr TileVariant t = kF64Variants[V];
    dispatch_tb(track, beads, [&](auto tr, auto bd) {
        hipLaunchKernelGGL(
            (k_step_f64_deferred<t.tx, t.ty, t.tz, deferred_k(t), decltype(tr)::value, decltype(bd)::value, RULE, SPHERE>),
            dim3(kDeferBlocks), dim3(kBlock), 0, s, A);
    });
}

template <int V>
static void launch_multi_v(const StepArgs *args, int nbox, bool track, bool beads, hipStream_t s)
{
    constexpr TileVariant t = kF64Variants[V];
    dispatch_tb(track, beads, [&](auto tr, auto bd) {
        hipLaunchKernelGGL(
            (k_step_f64_deferred_multi<t.tx, t.ty, t.tz, deferred_k(t), decltype(tr)::value, decltype(bd)::value>),
            dim3(kDeferBlocks, nbox), dim3(kBlock), 0, s, args);
    });
}

void launch_deferred_f64(const StepArgs &A, int variant, bool track, bool beads, hipStream_t s)
{
    if (A.rule == kRuleAt)
        launch_v<kAtVariant, kRuleAt, false>(A, track, beads, s);
    else if (A.rule == kRuleAtA)
        launch_v<kAtVariant, kRuleAtA, false>(A, track, beads, s);
    else if (variant == 1)
        A.sphere ? launch_v<1, kRuleSrd, true>(A, track, beads, s) : launch_v<1, kRuleSrd, false>(A, track, beads, s);
    else
        A.sphere ? launch_v<0, kRuleSrd, true>(A, track, beads, s) : launch_v<0, kRuleSrd, false>(A, track, beads, s);
}

void launch_deferred_multi_f64(const StepArgs *args, int nbox, int variant, bool track, bool beads, hipStream_t s)
{
    if (variant == 1)
        launch_multi_v<1>(args, nbox, track, beads, s);
    else
        launch_multi_v<0>(args, nbox, track, beads, s);
}

}  // namespace mcm
