// The later passes of the FP32-bulk step: the bead pass (tiles holding
// beads, normal capacity) and the large pass (tiles above the capacity,
// deferred_k particles per thread).  A fixed number of workgroups per box
// stride over the box's list; the tile code is in mpcd_tile.hpp.
#include "mpcd_tile.hpp"

namespace mcm {

template <int TX, int TY, int TZ, int KP, bool TRACK, int PASS, int RULE = kRuleSrd, int OBJ = kObjNone>
__device__ __forceinline__ void f32_later_tiles(const StepArgs &A)
{
    const int n = PASS == kPassBeads ? A.diag->bead_n : min(A.diag->defer_n, kDeferCap);
    for (int i = blockIdx.x; i < n; i += gridDim.x) {
        const int tile = PASS == kPassBeads ? A.bead_tiles[i] : A.diag->defer_tile[i];
        step_f32_tile<TX, TY, TZ, KP, TRACK, true, PASS, RULE, OBJ>(A, tile);
        __syncthreads();
    }
}

template <int TX, int TY, int TZ, int KP, bool TRACK, int PASS, int RULE, int OBJ = kObjNone>
__global__ __launch_bounds__(kBlock) void k_step_f32_later(StepArgs A)
{
    f32_later_tiles<TX, TY, TZ, KP, TRACK, PASS, RULE, OBJ>(A);
}

// blockIdx.y = box
template <int TX, int TY, int TZ, int KP, bool TRACK, int PASS, int RULE = kRuleSrd, int OBJ = kObjNone>
__global__ __launch_bounds__(kBlock) void k_step_f32_later_multi(const StepArgs *__restrict__ args)
{
    f32_later_tiles<TX, TY, TZ, KP, TRACK, PASS, RULE, OBJ>(args[blockIdx.y]);
}

template <int V, int PASS, int RULE = kRuleSrd, int OBJ = kObjNone>
static void launch_pass(const StepArgs *A, const StepArgs *args, int nbox, bool track, hipStream_t s)
{
    constexpr TileVariant t = kF32Variants[V];
    constexpr int KP = PASS == kPassBeads ? t.k : deferred_k(t);
    const int gx = PASS == kPassBeads ? (A ? std::min(A->g.ntiles, kBeadPassBlocks) : kEnsBeadBlocks) : kDeferBlocks;
    if (A) {
        if (track)
            hipLaunchKernelGGL((k_step_f32_later<t.tx, t.ty, t.tz, KP, true, PASS, RULE, OBJ>), dim3(gx),
                               dim3(kBlock), 0, s, *A);
        else
            hipLaunchKernelGGL((k_step_f32_later<t.tx, t.ty, t.tz, KP, false, PASS, RULE, OBJ>), dim3(gx),
                               dim3(kBlock), 0, s, *A);
    } else if constexpr (OBJ != kObjWalls) {
        if (track)
            hipLaunchKernelGGL((k_step_f32_later_multi<t.tx, t.ty, t.tz, KP, true, PASS, RULE, OBJ>), dim3(gx, nbox),
                               dim3(kBlock), 0, s, args);
        else
            hipLaunchKernelGGL((k_step_f32_later_multi<t.tx, t.ty, t.tz, KP, false, PASS, RULE, OBJ>), dim3(gx, nbox),
                               dim3(kBlock), 0, s, args);
    }
}

template <int V>
static void launch_v(const StepArgs *A, const StepArgs *args, int nbox, bool track, int pass, hipStream_t s)
{
    if (pass == kPassBeads)
        launch_pass<V, kPassBeads>(A, args, nbox, track, s);
    else
        launch_pass<V, kPassLarge>(A, args, nbox, track, s);
}

static void launch(const StepArgs *A, const StepArgs *args, int nbox, int variant, bool track, int pass,
                   hipStream_t s)
{
    switch (variant) {
    case 1: launch_v<1>(A, args, nbox, track, pass, s); break;
    case 2: launch_v<2>(A, args, nbox, track, pass, s); break;
    default: launch_v<0>(A, args, nbox, track, pass, s); break;
    }
}

// MPC-AT: one box (A), or every box of an ensemble (args, nbox)
template <int RULE>
static void launch_at(const StepArgs *A, const StepArgs *args, int nbox, bool track, int pass, hipStream_t s)
{
    if (pass == kPassBeads)
        launch_pass<kAtVariant, kPassBeads, RULE>(A, args, nbox, track, s);
    else
        launch_pass<kAtVariant, kPassLarge, RULE>(A, args, nbox, track, s);
}

void launch_later_f32(const StepArgs &A, int variant, bool track, int pass, hipStream_t s)
{
    if (A.sphere) {  // variant 0 (the context's with a sphere or walls)
        if (pass == kPassBeads)
            launch_pass<0, kPassBeads, kRuleSrd, kObjSphere>(&A, nullptr, 1, track, s);
        else
            launch_pass<0, kPassLarge, kRuleSrd, kObjSphere>(&A, nullptr, 1, track, s);
        return;
    }
    if (A.walls) {
        if (pass == kPassBeads)
            launch_pass<0, kPassBeads, kRuleSrd, kObjWalls>(&A, nullptr, 1, track, s);
        else
            launch_pass<0, kPassLarge, kRuleSrd, kObjWalls>(&A, nullptr, 1, track, s);
        return;
    }
    if (A.rule == kRuleAt)
        launch_at<kRuleAt>(&A, nullptr, 1, track, pass, s);
    else if (A.rule == kRuleAtA)
        launch_at<kRuleAtA>(&A, nullptr, 1, track, pass, s);
    else
        launch(&A, nullptr, 1, variant, track, pass, s);
}

void launch_later_multi(bool f32, const StepArgs *args, int nbox, int variant, bool track, int pass, hipStream_t s,
                        int rule, bool sphere)
{
    if (!f32)
        launch_later_multi_f64(args, nbox, variant, track, pass, s, rule, sphere);
    else if (sphere && pass == kPassBeads)
        launch_pass<0, kPassBeads, kRuleSrd, kObjSphere>(nullptr, args, nbox, track, s);
    else if (sphere)
        launch_pass<0, kPassLarge, kRuleSrd, kObjSphere>(nullptr, args, nbox, track, s);
    else if (rule == kRuleAt)
        launch_at<kRuleAt>(nullptr, args, nbox, track, pass, s);
    else if (rule == kRuleAtA)
        launch_at<kRuleAtA>(nullptr, args, nbox, track, pass, s);
    else
        launch(nullptr, args, nbox, variant, track, pass, s);
}

}  // namespace mcm
