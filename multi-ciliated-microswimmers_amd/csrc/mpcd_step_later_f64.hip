// The later passes of the FP64 step: the bead pass (tiles holding beads,
// normal capacity) and the large pass (tiles above the capacity,
// deferred_k particles per thread).  A fixed number of workgroups per box
// stride over the box's list; the tile code is in mpcd_tile.hpp.
#include "mpcd_tile.hpp"

namespace mcm {

template <int TX, int TY, int TZ, int KP, bool TRACK, int RULE, bool SPHERE, int PASS>
__device__ __forceinline__ void f64_later_tiles(const StepArgs &A)
{
    const int n = PASS == kPassBeads ? A.diag->bead_n : min(A.diag->defer_n, kDeferCap);
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int tile = PASS == kPassBeads ? A.bead_tiles[i] : A.diag->defer_tile[i];
        step_f64_tile<TX, TY, TZ, KP, TRACK, true, RULE, SPHERE, PASS>(A, tile);
        __syncthreads();
    }
}

template <int TX, int TY, int TZ, int KP, bool TRACK, int RULE, bool SPHERE, int PASS>
__global__ __launch_bounds__(kBlock) void k_step_f64_later(StepArgs A)
{
    f64_later_tiles<TX, TY, TZ, KP, TRACK, RULE, SPHERE, PASS>(A);
}

// blockIdx.y = box
template <int TX, int TY, int TZ, int KP, bool TRACK, int PASS, int RULE = kRuleSrd, bool SPHERE = false>
__global__ __launch_bounds__(kBlock) void k_step_f64_later_multi(const StepArgs *__restrict__ args)
{
    f64_later_tiles<TX, TY, TZ, KP, TRACK, RULE, SPHERE, PASS>(args[blockIdx.y]);
}

template <int V, int RULE, bool SPHERE, int PASS>
static void launch_pass(const StepArgs &A, bool track, hipStream_t s)
{
    constexpr TileVariant t = kF64Variants[V];
    constexpr int KP = PASS == kPassBeads ? t.k : deferred_k(t);
    const int gx = PASS == kPassBeads ? std::min(A.g.ntiles, kBeadPassBlocks) : kDeferBlocks;
    if (track)
        hipLaunchKernelGGL((k_step_f64_later<t.tx, t.ty, t.tz, KP, true, RULE, SPHERE, PASS>), dim3(gx), dim3(kBlock),
                           0, s, A);
    else
        hipLaunchKernelGGL((k_step_f64_later<t.tx, t.ty, t.tz, KP, false, RULE, SPHERE, PASS>), dim3(gx),
                           dim3(kBlock), 0, s, A);
}

template <int V, int RULE, bool SPHERE>
static void launch_v(const StepArgs &A, bool track, int pass, hipStream_t s)
{
    if (pass == kPassBeads)
        launch_pass<V, RULE, SPHERE, kPassBeads>(A, track, s);
    else
        launch_pass<V, RULE, SPHERE, kPassLarge>(A, track, s);
}

template <int V, int PASS, int RULE = kRuleSrd, bool SPHERE = false>
static void launch_multi_pass(const StepArgs *args, int nbox, bool track, hipStream_t s)
{
    constexpr TileVariant t = kF64Variants[V];
    constexpr int KP = PASS == kPassBeads ? t.k : deferred_k(t);
    const int gx = PASS == kPassBeads ? kEnsBeadBlocks : kDeferBlocks;
    if (track)
        hipLaunchKernelGGL((k_step_f64_later_multi<t.tx, t.ty, t.tz, KP, true, PASS, RULE, SPHERE>), dim3(gx, nbox),
                           dim3(kBlock), 0, s, args);
    else
        hipLaunchKernelGGL((k_step_f64_later_multi<t.tx, t.ty, t.tz, KP, false, PASS, RULE, SPHERE>), dim3(gx, nbox),
                           dim3(kBlock), 0, s, args);
}

template <int V>
static void launch_multi_sphere(const StepArgs *args, int nbox, bool track, int pass, hipStream_t s)
{
    if (pass == kPassBeads)
        launch_multi_pass<V, kPassBeads, kRuleSrd, true>(args, nbox, track, s);
    else
        launch_multi_pass<V, kPassLarge, kRuleSrd, true>(args, nbox, track, s);
}

template <int RULE>
static void launch_multi_at(const StepArgs *args, int nbox, bool track, int pass, hipStream_t s)
{
    if (pass == kPassBeads)
        launch_multi_pass<kAtVariant, kPassBeads, RULE>(args, nbox, track, s);
    else
        launch_multi_pass<kAtVariant, kPassLarge, RULE>(args, nbox, track, s);
}

void launch_later_f64(const StepArgs &A, int variant, bool track, int pass, hipStream_t s)
{
    if (A.rule == kRuleAt)
        launch_v<kAtVariant, kRuleAt, false>(A, track, pass, s);
    else if (A.rule == kRuleAtA)
        launch_v<kAtVariant, kRuleAtA, false>(A, track, pass, s);
    else if (variant == 1)
        A.sphere ? launch_v<1, kRuleSrd, true>(A, track, pass, s) : launch_v<1, kRuleSrd, false>(A, track, pass, s);
    else
        A.sphere ? launch_v<0, kRuleSrd, true>(A, track, pass, s) : launch_v<0, kRuleSrd, false>(A, track, pass, s);
}

void launch_later_multi_f64(const StepArgs *args, int nbox, int variant, bool track, int pass, hipStream_t s,
                            int rule, bool sphere)
{
    if (sphere && variant == 1)
        launch_multi_sphere<1>(args, nbox, track, pass, s);
    else if (sphere)
        launch_multi_sphere<0>(args, nbox, track, pass, s);
    else if (rule == kRuleAt)
        launch_multi_at<kRuleAt>(args, nbox, track, pass, s);
    else if (rule == kRuleAtA)
        launch_multi_at<kRuleAtA>(args, nbox, track, pass, s);
    else if (variant == 1)
        pass == kPassBeads ? launch_multi_pass<1, kPassBeads>(args, nbox, track, s)
                           : launch_multi_pass<1, kPassLarge>(args, nbox, track, s);
    else
        pass == kPassBeads ? launch_multi_pass<0, kPassBeads>(args, nbox, track, s)
                           : launch_multi_pass<0, kPassLarge>(args, nbox, track, s);
}

}  // namespace mcm
