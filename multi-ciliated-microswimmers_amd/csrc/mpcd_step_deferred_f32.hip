// The deferred pass of the FP32-bulk step: tiles above the step kernel's
// capacity, stepped by the same tile code with the larger capacity
// deferred_k (kDeferBlocks workgroups per box, striding over the box's
// deferred list).
#include "mpcd_tile.hpp"

namespace mcm {

template <int TX, int TY, int TZ, int KB, bool TRACK, bool BEADS>
__device__ __forceinline__ void f32_deferred_tiles(const StepArgs &A)
{
    const int n = min(A.diag->defer_n, kDeferCap);
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        step_f32_tile<TX, TY, TZ, KB, TRACK, BEADS, true>(A, A.diag->defer_tile[i]);
        __syncthreads();
    }
}

template <int TX, int TY, int TZ, int KB, bool TRACK, bool BEADS>
__global__ __launch_bounds__(kBlock) void k_step_f32_deferred(StepArgs A)
{
    f32_deferred_tiles<TX, TY, TZ, KB, TRACK, BEADS>(A);
}

// blockIdx.y = box
template <int TX, int TY, int TZ, int KB, bool TRACK, bool BEADS>
__global__ __launch_bounds__(kBlock) void k_step_f32_deferred_multi(const StepArgs *__restrict__ args)
{
    f32_deferred_tiles<TX, TY, TZ, KB, TRACK, BEADS>(args[blockIdx.y]);
}

template <int V>
static void launch_v(const StepArgs *A, const StepArgs *args, int nbox, bool track, bool beads, hipStream_t s)
{
    constexpr TileVariant t = kF32Variants[V];
    dispatch_tb(track, beads, [&](auto tr, auto bd) {
        constexpr bool TR = decltype(tr)::value, BD = decltype(bd)::value;
        if (A)
            hipLaunchKernelGGL((k_step_f32_deferred<t.tx, t.ty, t.tz, deferred_k(t), TR, BD>), dim3(kDeferBlocks),
                               dim3(kBlock), 0, s, *A);
        else
            hipLaunchKernelGGL((k_step_f32_deferred_multi<t.tx, t.ty, t.tz, deferred_k(t), TR, BD>),
                               dim3(kDeferBlocks, nbox), dim3(kBlock), 0, s, args);
    });
}

static void launch(const StepArgs *A, const StepArgs *args, int nbox, int variant, bool track, bool beads,
                   hipStream_t s)
{
    switch (variant) {
    case 1: launch_v<1>(A, args, nbox, track, beads, s); break;
    case 2: launch_v<2>(A, args, nbox, track, beads, s); break;
    default: launch_v<0>(A, args, nbox, track, beads, s); break;
    }
}

void launch_deferred_f32(const StepArgs &A, int variant, bool track, bool beads, hipStream_t s)
{
    launch(&A, nullptr, 1, variant, track, beads, s);
}

void launch_deferred_multi(bool f32, const StepArgs *args, int nbox, int variant, bool track, bool beads,
                           hipStream_t s)
{
    if (f32)
        launch(nullptr, args, nbox, variant, track, beads, s);
    else
        launch_deferred_multi_f64(args, nbox, variant, track, beads, s);
}

}  // namespace mcm
