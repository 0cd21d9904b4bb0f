// C ABI of the MPCD step (include/mpcd_abi.h): context, buffers, launches.
#include "../../include/mpcd_abi.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "mpcd_aux.hpp"
#include "mpcd_md.hpp"

using namespace mcm;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg)
{
    g_err = msg;
    return code;
}

#define HIPCHK(expr)                                                                             \
    do {                                                                                         \
        hipError_t e_ = (expr);                                                                  \
        if (e_ != hipSuccess)                                                                    \
            return fail(MPCD_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

// Host Philox4x64-10, identical to the device one (shift draws).
void philox_host(const uint64_t c_in[4], uint64_t k0, uint64_t k1, uint64_t out[4])
{
    uint64_t c0 = c_in[0], c1 = c_in[1], c2 = c_in[2], c3 = c_in[3];
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B97F4A7C15ull;
            k1 += 0xBB67AE8584CAA73Bull;
        }
        const unsigned __int128 p0 = (unsigned __int128)0xD2E7470EE14C6C93ull * c0;
        const unsigned __int128 p1 = (unsigned __int128)0xCA5A826395121157ull * c2;
        const uint64_t n0 = (uint64_t)(p1 >> 64) ^ c1 ^ k0, n2 = (uint64_t)(p0 >> 64) ^ c3 ^ k1;
        c1 = (uint64_t)p1;
        c3 = (uint64_t)p0;
        c0 = n0;
        c2 = n2;
    }
    out[0] = c0;
    out[1] = c1;
    out[2] = c2;
    out[3] = c3;
}

// Random grid shift of a step on the 2^-24 grid inside (0,1)^3: the
// counter-based replacement of MPCD::shift_boxes (mpcd.cpp:249-253).
void shift_of_step(uint64_t seed, uint64_t step, double a[3])
{
    const uint64_t ctr[4] = {0, step, 0, 0x53484654ull};
    uint64_t o[4];
    philox_host(ctr, seed, kKeyHi, o);
    for (int k = 0; k < 3; ++k)
        a[k] = (double)(1 + (o[k] >> 40) % 16777215ull) * (1.0 / 16777216.0);
}

template <class T>
int dalloc(T **p, size_t count)
{
    *p = nullptr;
    if (count == 0)
        count = 1;
    hipError_t e = hipMalloc((void **)p, count * sizeof(T));
    if (e != hipSuccess)
        return fail(MPCD_ERR_HIP, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes: " + hipGetErrorString(e));
    return MPCD_OK;
}

}  // namespace

struct mpcd_ctx {
    mpcd_params p{};
    Geom g{};
    bool f32 = false;
    bool track = false;
    int variant = 0;
    int W = 0;
    int nc = 0;  // cells per tile of the step kernel
    uint32_t *home[2] = {nullptr, nullptr};
    uint32_t *ob[2] = {nullptr, nullptr};
    int *home_cnt[2] = {nullptr, nullptr};
    int *ob_cnt[2] = {nullptr, nullptr};
    uint32_t *ovf = nullptr;
    int *ovf_cnt = nullptr;  // two counters, one per step parity
    int *flags = nullptr;
    Diag *diag = nullptr;
    double *sum3 = nullptr;
    int cur = 0;
    bool loaded = false;
    double a[3] = {0, 0, 0};
    uint64_t step = 0;
    long long n = 0;
    hipStream_t stream = nullptr;
    double *inj_gamma = nullptr;
    double *inj_axis = nullptr;
    double *inj_virt = nullptr;
    double *inj_ran = nullptr;
    size_t inj_ran_cap = 0;
    double *inj_sph = nullptr;
    // rigid sphere (device state and impulse accumulators)
    SphereDev *sphere = nullptr;
    // beads
    int nb = 0, nb_cap = 0, nb_last = 0;
    double *b_pos3 = nullptr, *b_vel4_raw = nullptr, *b_vel4 = nullptr, *b_out4 = nullptr;
    unsigned *b_keys = nullptr, *b_vals = nullptr, *b_skeys = nullptr, *b_svals = nullptr;
    int *b_lc = nullptr, *b_idx = nullptr, *b_tile_start = nullptr;
    int *bead_tiles = nullptr;  // the bead pass's tile list (one entry per tile at most)
    void *b_temp = nullptr;
    size_t b_temp_bytes = 0;
    // attached swimmer
    MdSystem *md = nullptr;
    // the ensemble this context is stepped by (state changes are refused)
    mpcd_ensemble *ens = nullptr;
    // debug
    int *dbg_cell = nullptr, *dbg_rank = nullptr;
    long long dbg_n = 0;
    // timing
    bool timing = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    double timed_ms = 0.0;
    long long timed_launches = 0;
};

struct mpcd_ensemble {
    std::vector<mpcd_ctx *> box;
    std::vector<hipStream_t> own_stream;  // the members' streams, restored at destroy
    hipStream_t stream = nullptr;         // shared by every member while in the ensemble
    bool f32 = false, track = false, beads = false;
    int variant = 0;
    int rule = kRuleSrd;     // the members' collision rule
    bool sphere = false;     // every member has a rigid sphere
    int ntiles = 0;
    int nsw = 0;
    EnsBox *d_boxes = nullptr;
    StepArgs *d_args = nullptr;
    int *d_tile_start = nullptr;
    MdBox *d_md = nullptr;
    int *d_mdblk = nullptr;  // the wide MD path's workgroup tables
    MdWide wide{};
    bool md_wide = true;     // MD: one launch per phase over all swimmers (else a workgroup per swimmer)
    int md_block = kMdWideBlock;  // threads per workgroup of the wide MD launches
    double tstep = -1.0;     // the step length the device tables and graphs are for
    bool use_graph = true;
    // a captured step per union of the swimmers' refresh schedules (the wide
    // MD path launches the refresh only at the evaluations where some
    // swimmer refreshes)
    SchedGraphs graphs;
};

namespace {

int check_flags(mpcd_ctx *c)
{
    int f = 0;
    HIPCHK(hipMemcpyAsync(&f, c->flags, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (f) {
        std::string m = "device capacity/state error flags=" + std::to_string(f) + ":";
        if (f & kFlagTileCap) m += " tile particle capacity exceeded;";
        if (f & kFlagOverflowCap) m += " overflow list full;";
        if (f & kFlagHomeCap) m += " home segment full;";
        if (f & kFlagVelocity) m += " FP32-bulk velocity or displacement out of range;";
        if (f & kFlagMisbinned) m += " particle outside its tile;";
        return fail(MPCD_ERR_CAPACITY, m);
    }
    return MPCD_OK;
}

void free_all(mpcd_ctx *c)
{
    auto F = [](void *p) {
        if (p)
            (void)hipFree(p);
    };
    for (int b = 0; b < 2; ++b) {
        F(c->home[b]);
        F(c->ob[b]);
        F(c->home_cnt[b]);
        F(c->ob_cnt[b]);
    }
    F(c->ovf);
    F(c->ovf_cnt);
    F(c->flags);
    F(c->diag);
    F(c->sum3);
    F(c->inj_gamma);
    F(c->inj_virt);
    F(c->inj_ran);
    F(c->inj_sph);
    F(c->sphere);
    F(c->inj_axis);
    F(c->b_pos3);
    F(c->b_vel4_raw);
    F(c->b_vel4);
    F(c->b_out4);
    F(c->b_keys);
    F(c->b_vals);
    F(c->b_skeys);
    F(c->b_svals);
    F(c->b_lc);
    F(c->b_idx);
    F(c->b_tile_start);
    F(c->bead_tiles);
    F(c->b_temp);
    F(c->dbg_cell);
    F(c->dbg_rank);
    if (c->md) {
        md_free(*c->md);
        delete c->md;
        c->md = nullptr;
    }
    for (auto &e : c->events) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
}

int harvest_events(mpcd_ctx *c)
{
    if (c->events.empty())
        return MPCD_OK;
    HIPCHK(hipStreamSynchronize(c->stream));
    for (auto &e : c->events) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, e.first, e.second));
        c->timed_ms += ms;
        c->timed_launches += 1;
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    c->events.clear();
    return MPCD_OK;
}

int refuse_in_ensemble(const mpcd_ctx *c)
{
    return c->ens ? fail(MPCD_ERR_STATE, "the context is stepped by an ensemble: destroy the ensemble first") : MPCD_OK;
}

// The step arguments that do not change from step to step, for the buffer
// parity cur (input buffers cur, output buffers 1 - cur).  The caller adds
// the shift, the step counter, the overflow counter of the step's parity,
// injected draws and beads.
StepArgs base_args(const mpcd_ctx *c, int cur, double tstep)
{
    StepArgs A{};
    A.g = c->g;
    const int in = cur, out = 1 - cur;
    A.in_home = c->home[in];
    A.in_ob = c->ob[in];
    A.in_home_cnt = c->home_cnt[in];
    A.in_ob_cnt = c->ob_cnt[in];
    A.out_home = c->home[out];
    A.out_ob = c->ob[out];
    A.out_home_cnt = c->home_cnt[out];
    A.out_ob_cnt = c->ob_cnt[out];
    A.ovf = c->ovf;
    A.flags = c->flags;
    A.h = tstep;
    A.hf = (float)tstep;
    A.calpha = c->p.calpha;
    A.salpha = c->p.salpha;
    A.kT = c->p.kT;
    A.thermostat = c->p.thermostat;
    A.seed = c->p.seed;
    A.walls = c->p.walls;
    A.rho = c->p.rho;
    A.vw[0] = c->p.wall_v[0];
    A.vw[1] = c->p.wall_v[1];
    A.rule = c->p.collision;
    A.nf_total = c->n;
    A.sphere = c->sphere;
    A.diag = c->diag;
    A.bead_tiles = c->bead_tiles;
    A.dbg_cell = c->dbg_cell;
    A.dbg_rank = c->dbg_rank;
    return A;
}

// Bead arrays for at least nb beads.
int ensure_bead_cap(mpcd_ctx *ctx, int nb)
{
    if (nb <= ctx->nb_cap)
        return MPCD_OK;
    const int cap = std::max(nb, 1024);
    auto F = [](void *p) {
        if (p)
            (void)hipFree(p);
    };
    F(ctx->b_pos3); F(ctx->b_vel4_raw); F(ctx->b_vel4); F(ctx->b_out4);
    F(ctx->b_keys); F(ctx->b_vals); F(ctx->b_skeys); F(ctx->b_svals);
    F(ctx->b_lc); F(ctx->b_idx); F(ctx->b_temp);
    ctx->nb_cap = 0;
    int rc;
    if ((rc = dalloc(&ctx->b_pos3, (size_t)cap * 3)) || (rc = dalloc(&ctx->b_vel4_raw, (size_t)cap * 4)) ||
        (rc = dalloc(&ctx->b_vel4, (size_t)cap * 4)) || (rc = dalloc(&ctx->b_out4, (size_t)cap * 4)) ||
        (rc = dalloc(&ctx->b_keys, (size_t)cap)) || (rc = dalloc(&ctx->b_vals, (size_t)cap)) ||
        (rc = dalloc(&ctx->b_skeys, (size_t)cap)) || (rc = dalloc(&ctx->b_svals, (size_t)cap)) ||
        (rc = dalloc(&ctx->b_lc, (size_t)cap)) || (rc = dalloc(&ctx->b_idx, (size_t)cap)))
        return rc;
    ctx->b_temp_bytes = bead_sort_temp_bytes(cap);
    if ((rc = dalloc((char **)&ctx->b_temp, ctx->b_temp_bytes)))
        return rc;
    ctx->nb_cap = cap;
    return MPCD_OK;
}

}  // namespace

extern "C" {

const char *mpcd_version(void) { return "mpcd-gfx950 0.1"; }

const char *mpcd_last_error(void) { return g_err.c_str(); }

int mpcd_create(const mpcd_params *p, mpcd_ctx **out)
{
    if (!p || !out)
        return fail(MPCD_ERR_ARG, "null argument");
    *out = nullptr;
    for (int k = 0; k < 3; ++k)
        if (p->size[k] < 1 || p->size[k] > 1024)
            return fail(MPCD_ERR_ARG, "box size per axis must be in [1, 1024]");
    if (p->rho < 1 || p->rho > 10)
        return fail(MPCD_ERR_ARG, "rho must be in [1, 10] (tile capacity is sized for <= 10 per cell)");
    if (p->precision != MPCD_FP64 && p->precision != MPCD_FP32BULK)
        return fail(MPCD_ERR_ARG, "precision must be MPCD_FP64 or MPCD_FP32BULK");
    if (!(p->kT > 0))
        return fail(MPCD_ERR_ARG, "kT must be positive");
    if (p->walls < MPCD_WALLS_NONE || p->walls > MPCD_WALLS_SLIP)
        return fail(MPCD_ERR_ARG, "walls must be MPCD_WALLS_NONE, MPCD_WALLS_NOSLIP or MPCD_WALLS_SLIP");
    if (p->walls && p->precision == MPCD_FP32BULK && p->tile_variant > 0)
        return fail(MPCD_ERR_ARG, "the FP32-bulk form takes planar walls with tile variant 0 (the default)");
    if (p->walls && p->size[2] < 2)
        return fail(MPCD_ERR_ARG, "planar walls need at least 2 cells along z");
    if (p->collision < MPCD_SRD || p->collision > MPCD_AT_A)
        return fail(MPCD_ERR_ARG, "collision must be MPCD_SRD, MPCD_AT or MPCD_AT_A");
    if (p->collision != MPCD_SRD && p->walls)
        return fail(MPCD_ERR_ARG, "MPC-AT needs a box without walls");
    if (p->collision != MPCD_SRD && p->tile_variant != -1 && p->tile_variant != kAtVariant)
        return fail(MPCD_ERR_ARG, "MPC-AT runs on tile variant 1");
    mpcd_ctx *c = new mpcd_ctx();
    c->p = *p;
    c->f32 = p->precision == MPCD_FP32BULK;
    c->track = p->track_ids != 0;
    c->W = c->f32 ? (c->track ? 8 : 7) : (c->track ? 14 : 12);
    Geom &g = c->g;
    const int nvar = c->f32 ? kNumF32Variants : kNumF64Variants;
    if (p->tile_variant < -1 || p->tile_variant >= nvar) {
        delete c;
        return fail(MPCD_ERR_ARG, "tile_variant out of range");
    }
    c->variant = p->tile_variant >= 0 ? p->tile_variant
                                      : (p->collision != MPCD_SRD ? kAtVariant : (c->f32 ? 0 : kF64DefaultVariant));
    const TileVariant tv = c->f32 ? kF32Variants[c->variant] : kF64Variants[c->variant];
    const int T[3] = {tv.tx, tv.ty, tv.tz};
    const int K = tv.k;
    c->nc = T[0] * T[1] * T[2];
    // the step kernel holds 1.2 times the mean load (7 standard deviations of
    // a Poisson count at rho 10); denser tiles go to the large pass
    if (1.2 * p->rho * c->nc > K * kBlock) {
        delete c;
        return fail(MPCD_ERR_ARG, "rho too large for this tile variant's particle capacity");
    }
    for (int k = 0; k < 3; ++k) {
        g.L[k] = p->size[k];
        g.invL[k] = 1.0 / p->size[k];
        g.T[k] = T[k];
        g.nt[k] = (p->size[k] + T[k] - 1) / T[k];
    }
    g.ntiles = g.nt[0] * g.nt[1] * g.nt[2];
    g.cap_home = deferred_k(tv) * kBlock;  // a dense tile's records stay at home for the deferred pass
    const double rho = p->rho;
    int off = 0;
    for (int d = 0; d < 26; ++d) {
        const int i = d < 13 ? d : d + 1;
        const int v[3] = {i / 9 - 1, (i / 3) % 3 - 1, i % 3 - 1};
        const int nz = (v[0] != 0) + (v[1] != 0) + (v[2] != 0);
        int cap;
        if (nz == 1) {  // face: area normal to the moving axis
            int area = 1;
            for (int k = 0; k < 3; ++k)
                if (!v[k]) area *= T[k];
            cap = (int)std::ceil(1.6 * rho * area) + 64;
        } else if (nz == 2) {  // edge: length along the resting axis
            int len = 1;
            for (int k = 0; k < 3; ++k)
                if (!v[k]) len = T[k];
            cap = (int)std::ceil(2.5 * rho * len) + 32;
        } else {
            cap = (int)std::ceil(4.0 * rho) + 32;
        }
        g.ob_cap[d] = cap;
        g.ob_off[d] = off;
        off += cap;
    }
    g.ob_stride = off;
    const long long ncells = (long long)g.L[0] * g.L[1] * g.L[2];
    g.cap_ovf = (int)std::max<long long>(65536, ncells * p->rho / 100);

    int rc;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(MPCD_ERR_HIP, "hipStreamCreate failed");
    }
    const size_t home_words = (size_t)g.ntiles * g.cap_home * c->W;
    const size_t ob_words = (size_t)g.ntiles * g.ob_stride * c->W;
    for (int b = 0; b < 2; ++b) {
        if ((rc = dalloc(&c->home[b], home_words)) || (rc = dalloc(&c->ob[b], ob_words)) ||
            (rc = dalloc(&c->home_cnt[b], (size_t)g.ntiles)) || (rc = dalloc(&c->ob_cnt[b], (size_t)g.ntiles * 26))) {
            free_all(c);
            delete c;
            return rc;
        }
    }
    if ((rc = dalloc(&c->ovf, (size_t)g.cap_ovf * c->W)) || (rc = dalloc(&c->ovf_cnt, 2)) || (rc = dalloc(&c->flags, 1)) ||
        (rc = dalloc(&c->diag, 1)) || (rc = dalloc(&c->sum3, (size_t)kCreateSum3)) ||
        (rc = dalloc(&c->b_tile_start, (size_t)g.ntiles + 1)) || (rc = dalloc(&c->bead_tiles, (size_t)g.ntiles))) {
        free_all(c);
        delete c;
        return rc;
    }
    for (int b = 0; b < 2; ++b) {
        fill_i32(c->home_cnt[b], 0, g.ntiles, c->stream);
        fill_i32(c->ob_cnt[b], 0, (long long)g.ntiles * 26, c->stream);
    }
    fill_i32(c->ovf_cnt, 0, 2, c->stream);
    fill_i32(c->flags, 0, 1, c->stream);
    fill_u64(c->diag, 0ull, (long long)(sizeof(Diag) / 8), c->stream);
    fill_i32(c->b_tile_start, 0, g.ntiles + 1, c->stream);
    shift_of_step(p->seed, 0, c->a);
    if (hipStreamSynchronize(c->stream) != hipSuccess) {
        free_all(c);
        delete c;
        return fail(MPCD_ERR_HIP, "initialisation failed");
    }
    *out = c;
    return MPCD_OK;
}

int mpcd_destroy(mpcd_ctx *ctx)
{
    if (!ctx)
        return MPCD_OK;
    if (int rc = refuse_in_ensemble(ctx))
        return rc;
    if (ctx->stream)
        (void)hipStreamSynchronize(ctx->stream);
    free_all(ctx);
    delete ctx;
    return MPCD_OK;
}

int mpcd_set_shift(mpcd_ctx *ctx, const double a[3])
{
    if (!ctx || !a)
        return fail(MPCD_ERR_ARG, "null argument");
    if (ctx->loaded)
        return fail(MPCD_ERR_STATE, "the shift can only be injected before the fluid is uploaded");
    for (int k = 0; k < 3; ++k)
        ctx->a[k] = a[k];
    return MPCD_OK;
}

int mpcd_get_shift(mpcd_ctx *ctx, double a[3])
{
    if (!ctx || !a)
        return fail(MPCD_ERR_ARG, "null argument");
    for (int k = 0; k < 3; ++k)
        a[k] = ctx->a[k];
    return MPCD_OK;
}

int mpcd_draw_shift(mpcd_ctx *ctx, double a[3])
{
    if (!ctx || !a)
        return fail(MPCD_ERR_ARG, "null argument");
    shift_of_step(ctx->p.seed, ctx->step + 1, a);
    return MPCD_OK;
}

static int upload_staged(mpcd_ctx *c, const double *dpos, const double *dvel, long long n)
{
    const Geom &g = c->g;
    fill_i32(c->home_cnt[c->cur], 0, g.ntiles, c->stream);
    fill_i32(c->ob_cnt[c->cur], 0, (long long)g.ntiles * 26, c->stream);
    launch_upload(g, c->f32, c->track, dpos, dvel, n, c->a, c->home[c->cur], c->home_cnt[c->cur], c->flags, c->stream);
    HIPCHK(hipGetLastError());
    c->n = n;
    c->loaded = true;
    return check_flags(c);
}

int mpcd_upload_fluid(mpcd_ctx *ctx, const double *pos3, const double *vel4, long long n)
{
    if (!ctx || (!pos3 && n) || (!vel4 && n) || n < 0)
        return fail(MPCD_ERR_ARG, "bad upload arguments");
    if (int rc = refuse_in_ensemble(ctx))
        return rc;
    if (ctx->track && n > 0x7fffffffLL)
        return fail(MPCD_ERR_ARG, "ids are 31-bit");
    for (long long i = 0; i < n; ++i)
        if (vel4[4 * i + 3] != 1.0)
            return fail(MPCD_ERR_ARG, "solvent mass column must be 1.0 (fluid.cpp:66)");
    double *dpos = nullptr, *dvel = nullptr;
    int rc;
    if ((rc = dalloc(&dpos, (size_t)n * 3)))
        return rc;
    if ((rc = dalloc(&dvel, (size_t)n * 4))) {
        (void)hipFree(dpos);
        return rc;
    }
    rc = MPCD_OK;
    if (n && (hipMemcpyAsync(dpos, pos3, (size_t)n * 3 * sizeof(double), hipMemcpyHostToDevice, ctx->stream) != hipSuccess ||
              hipMemcpyAsync(dvel, vel4, (size_t)n * 4 * sizeof(double), hipMemcpyHostToDevice, ctx->stream) != hipSuccess))
        rc = fail(MPCD_ERR_HIP, "upload copy failed");
    if (!rc)
        rc = upload_staged(ctx, dpos, dvel, n);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(dpos);
    (void)hipFree(dvel);
    return rc;
}

int mpcd_create_fluid(mpcd_ctx *ctx, long long n)
{
    if (!ctx || n < 0)
        return fail(MPCD_ERR_ARG, "bad arguments");
    if (int rc = refuse_in_ensemble(ctx))
        return rc;
    double *dpos = nullptr, *dvel = nullptr;
    int rc;
    if ((rc = dalloc(&dpos, (size_t)n * 3)))
        return rc;
    if ((rc = dalloc(&dvel, (size_t)n * 4))) {
        (void)hipFree(dpos);
        return rc;
    }
    // positions over the whole box, walls or not: the reference's SLIP /
    // NOSLIP build creates z over [0, Lz) (fluid.cpp:59-62; the drop at
    // fluid.cpp:67-72 is compiled only with BOUNDS, which that build cannot
    // have, mpcd.cpp:389-392), and the first move bounces the particles
    // above the top wall z = Lz - 1 back in
    const double ext[3] = {(double)ctx->g.L[0], (double)ctx->g.L[1], (double)ctx->g.L[2]};
    launch_create_fluid(ext, ctx->p.kT, ctx->p.seed, n, dpos, dvel, ctx->sum3, ctx->stream);
    rc = upload_staged(ctx, dpos, dvel, n);
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(dpos);
    (void)hipFree(dvel);
    return rc;
}

// ---- fluid creation that allocates nothing: its scratch (the blocks'
// partial velocity sums and their total, kCreateSum3 doubles = 49,184 B) is
// this file-scope array, one copy per GPU for the life of the process.  In
// host code the array's name is no device address: the device address
// comes from hipGetSymbolAddress, once per device, and only that is used.
static __device__ __align__(256) double scratch[kCreateSum3];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};

int scratch_address(double **out)
{
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        return fail(MPCD_ERR_ARG, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(MPCD_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[dev] = static_cast<double *>(p);
    }
    *out = g_scratch[dev];
    return MPCD_OK;
}
}  // namespace

int mpcd_create_fluid_noalloc(mpcd_ctx *ctx, long long n)
{
    if (!ctx || n < 0)
        return fail(MPCD_ERR_ARG, "bad arguments");
    if (int rc = refuse_in_ensemble(ctx))
        return rc;
    if (ctx->track && n > 0x7fffffffLL)
        return fail(MPCD_ERR_ARG, "ids are 31-bit");
    const Geom &g = ctx->g;
    const long long most = (long long)g.cap_home * g.ntiles;
    if (n > most)
        return fail(MPCD_ERR_ARG, "n = " + std::to_string(n) + " exceeds the " + std::to_string(most) +
                                      " records the home segments hold");
    double *scr = nullptr;
    if (int rc = scratch_address(&scr))
        return rc;
    fill_i32(ctx->home_cnt[ctx->cur], 0, g.ntiles, ctx->stream);
    fill_i32(ctx->ob_cnt[ctx->cur], 0, (long long)g.ntiles * 26, ctx->stream);
    const double ext[3] = {(double)g.L[0], (double)g.L[1], (double)g.L[2]};
    launch_create_records(g, ctx->f32, ctx->track, ext, ctx->p.kT, ctx->p.seed, n, scr, ctx->a, ctx->home[ctx->cur],
                          ctx->home_cnt[ctx->cur], ctx->flags, ctx->stream);
    HIPCHK(hipGetLastError());
    ctx->n = n;
    ctx->loaded = true;
    return check_flags(ctx);
}

long long mpcd_count(mpcd_ctx *ctx) { return ctx ? ctx->n : -1; }

int mpcd_download_fluid(mpcd_ctx *ctx, double *pos3, double *vel4, long long n)
{
    if (!ctx || n != ctx->n || (n && (!pos3 || !vel4)))
        return fail(MPCD_ERR_ARG, "download size must equal the particle count");
    int rc = check_flags(ctx);
    if (rc)
        return rc;
    const Geom &g = ctx->g;
    int *tot = nullptr;
    long long *prefix = nullptr;
    double *dpos = nullptr, *dvel = nullptr;
    if ((rc = dalloc(&tot, (size_t)g.ntiles)) || (rc = dalloc(&prefix, (size_t)g.ntiles)) ||
        (rc = dalloc(&dpos, (size_t)n * 3)) || (rc = dalloc(&dvel, (size_t)n * 4))) {
        (void)hipFree(tot);
        (void)hipFree(prefix);
        (void)hipFree(dpos);
        return rc;
    }
    std::vector<int> htot(g.ntiles);
    std::vector<long long> hpre(g.ntiles);
    launch_tile_totals(g, ctx->home_cnt[ctx->cur], ctx->ob_cnt[ctx->cur], tot, ctx->stream);
    rc = MPCD_OK;
    if (hipMemcpyAsync(htot.data(), tot, sizeof(int) * g.ntiles, hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
        hipStreamSynchronize(ctx->stream) != hipSuccess)
        rc = fail(MPCD_ERR_HIP, "download totals failed");
    long long run = 0;
    for (int t = 0; t < g.ntiles && !rc; ++t) {
        hpre[t] = run;
        run += htot[t];
    }
    if (!rc && run != n)
        rc = fail(MPCD_ERR_STATE, "device holds " + std::to_string(run) + " particles, expected " + std::to_string(n));
    if (!rc) {
        if (hipMemcpyAsync(prefix, hpre.data(), sizeof(long long) * g.ntiles, hipMemcpyHostToDevice, ctx->stream) != hipSuccess)
            rc = fail(MPCD_ERR_HIP, "prefix copy failed");
    }
    if (!rc) {
        launch_download(g, ctx->f32, ctx->track, ctx->home[ctx->cur], ctx->ob[ctx->cur], ctx->home_cnt[ctx->cur],
                        ctx->ob_cnt[ctx->cur], prefix, ctx->a, dpos, dvel, ctx->stream);
        if (n && (hipMemcpyAsync(pos3, dpos, (size_t)n * 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess ||
                  hipMemcpyAsync(vel4, dvel, (size_t)n * 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream) != hipSuccess))
            rc = fail(MPCD_ERR_HIP, "download copy failed");
        if (hipStreamSynchronize(ctx->stream) != hipSuccess)
            rc = fail(MPCD_ERR_HIP, "download failed");
    }
    (void)hipFree(tot);
    (void)hipFree(prefix);
    (void)hipFree(dpos);
    (void)hipFree(dvel);
    return rc;
}

namespace {

// Copies beads in (from the host, or from an attached swimmer's device
// arrays) and bins them by (tile, local cell, index) under the current shift.
int stage_beads_impl(mpcd_ctx *ctx, const double *pos3, const double *vel4, int nb, hipMemcpyKind kind,
                     const unsigned char *mobile = nullptr)
{
    if (nb && ctx->f32 && (ctx->sphere || ctx->p.walls))
        return fail(MPCD_ERR_ARG, "the FP32-bulk form takes beads without a rigid sphere or walls");
    int rc = ensure_bead_cap(ctx, nb);
    if (rc)
        return rc;
    if (nb) {
        HIPCHK(hipMemcpyAsync(ctx->b_pos3, pos3, (size_t)nb * 3 * sizeof(double), kind, ctx->stream));
        HIPCHK(hipMemcpyAsync(ctx->b_vel4_raw, vel4, (size_t)nb * 4 * sizeof(double), kind, ctx->stream));
        HIPCHK(hipMemcpyAsync(ctx->b_out4, vel4, (size_t)nb * 4 * sizeof(double), kind, ctx->stream));
    }
    if (launch_bead_bin(ctx->g, ctx->nc, ctx->b_pos3, ctx->b_vel4_raw, mobile, nb, ctx->a, ctx->b_keys, ctx->b_vals,
                        ctx->b_skeys, ctx->b_svals, ctx->b_temp, ctx->b_temp_bytes, ctx->b_lc, ctx->b_idx,
                        ctx->b_vel4, ctx->b_tile_start, ctx->stream))
        return fail(MPCD_ERR_HIP, "bead sort failed");
    HIPCHK(hipGetLastError());
    ctx->nb = nb;
    return MPCD_OK;
}

}  // namespace

int mpcd_stage_beads(mpcd_ctx *ctx, const double *pos3, const double *vel4, int nb)
{
    if (!ctx || nb < 0 || (nb && (!pos3 || !vel4)))
        return fail(MPCD_ERR_ARG, "bad bead arguments");
    if (ctx->md)
        return fail(MPCD_ERR_STATE, "a swimmer is attached: its beads are binned on the device");
    if (int rc = refuse_in_ensemble(ctx))
        return rc;
    return stage_beads_impl(ctx, pos3, vel4, nb, hipMemcpyHostToDevice);
}

int mpcd_fetch_beads(mpcd_ctx *ctx, double *vel4, int nb)
{
    if (!ctx || nb != ctx->nb_last || (nb && !vel4))
        return fail(MPCD_ERR_ARG, "bead count must equal the beads of the last collision");
    if (nb)
        HIPCHK(hipMemcpyAsync(vel4, ctx->b_out4, (size_t)nb * 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MPCD_OK;
}

int mpcd_step(mpcd_ctx *ctx, double tstep, const double *next_shift, const double *inj_gamma, const double *inj_axis)
{
    if (ctx && ctx->p.walls && inj_gamma)
        return fail(MPCD_ERR_ARG, "with walls, inject through mpcd_step_injected (virtual-particle normals too)");
    if (ctx && ctx->p.collision != MPCD_SRD && inj_gamma)
        return fail(MPCD_ERR_ARG, "MPC-AT draws no Gamma variates or axes; inject through mpcd_step_injected");
    if (ctx && ctx->sphere && inj_gamma)
        return fail(MPCD_ERR_ARG, "with a sphere, inject through mpcd_step_injected (its normals too)");
    mpcd_draws d{};
    d.gamma = inj_gamma;
    d.axis = inj_axis;
    return mpcd_step_injected(ctx, tstep, next_shift, &d);
}

int mpcd_step_injected(mpcd_ctx *ctx, double tstep, const double *next_shift, const mpcd_draws *draws)
{
    if (!ctx)
        return fail(MPCD_ERR_ARG, "null context");
    if (int rc = refuse_in_ensemble(ctx))
        return rc;
    const mpcd_draws none{};
    const mpcd_draws &dr = draws ? *draws : none;
    const double *inj_gamma = dr.gamma, *inj_axis = dr.axis, *inj_virt = dr.virt, *inj_ran = dr.ran,
                 *inj_sph = dr.sph;
    if (ctx->sphere && (inj_gamma == nullptr) != (inj_sph == nullptr))
        return fail(MPCD_ERR_ARG, "with a sphere, inject gamma, axis and the sphere's normals together");
    if (!ctx->sphere && inj_sph)
        return fail(MPCD_ERR_ARG, "sphere normals given but no sphere is attached");
    if (!ctx->loaded)
        return fail(MPCD_ERR_STATE, "no fluid uploaded");
    if (ctx->p.collision != MPCD_SRD && (inj_gamma || inj_axis || inj_virt))
        return fail(MPCD_ERR_ARG, "MPC-AT takes injected normals per particle only (inj_ran)");
    if (ctx->p.collision == MPCD_SRD && inj_ran)
        return fail(MPCD_ERR_ARG, "inj_ran is for the MPC-AT rules");
    if (inj_ran && !ctx->track)
        return fail(MPCD_ERR_ARG, "injected MPC-AT normals are indexed by particle id: create with track_ids = 1");
    if ((inj_gamma == nullptr) != (inj_axis == nullptr))
        return fail(MPCD_ERR_ARG, "inject gamma and axis together");
    if (ctx->p.walls && (inj_gamma == nullptr) != (inj_virt == nullptr))
        return fail(MPCD_ERR_ARG, "with walls, inject gamma, axis and virtual-particle normals together");
    if (!ctx->p.walls && inj_virt)
        return fail(MPCD_ERR_ARG, "virtual-particle normals given but the box has no walls");
    const Geom &g = ctx->g;
    const long long ncells = (long long)g.L[0] * g.L[1] * g.L[2];
    int rc;
    if (inj_gamma) {
        if (!ctx->inj_gamma && ((rc = dalloc(&ctx->inj_gamma, (size_t)ncells)) || (rc = dalloc(&ctx->inj_axis, (size_t)ncells * 3))))
            return rc;
        HIPCHK(hipMemcpyAsync(ctx->inj_gamma, inj_gamma, sizeof(double) * ncells, hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(ctx->inj_axis, inj_axis, sizeof(double) * 3 * ncells, hipMemcpyHostToDevice, ctx->stream));
    }
    if (inj_virt) {
        if (!ctx->inj_virt && (rc = dalloc(&ctx->inj_virt, (size_t)ncells * 3)))
            return rc;
        HIPCHK(hipMemcpyAsync(ctx->inj_virt, inj_virt, sizeof(double) * 3 * ncells, hipMemcpyHostToDevice, ctx->stream));
    }
    if (inj_sph) {
        const size_t nsph = 3 * (size_t)ncells * (size_t)ctx->p.rho;
        if (!ctx->inj_sph && (rc = dalloc(&ctx->inj_sph, nsph)))
            return rc;
        HIPCHK(hipMemcpyAsync(ctx->inj_sph, inj_sph, sizeof(double) * nsph, hipMemcpyHostToDevice, ctx->stream));
    }
    const int nb_ran = (ctx->md && ctx->md->nb) ? ctx->md->nb : ctx->nb;
    const size_t nran = 3 * ((size_t)ctx->n + (size_t)nb_ran);
    if (inj_ran) {
        if (ctx->inj_ran_cap < nran) {
            if (ctx->inj_ran)
                (void)hipFree(ctx->inj_ran);
            ctx->inj_ran = nullptr;
            ctx->inj_ran_cap = 0;
            if ((rc = dalloc(&ctx->inj_ran, nran)))
                return rc;
            ctx->inj_ran_cap = nran;
        }
        HIPCHK(hipMemcpyAsync(ctx->inj_ran, inj_ran, sizeof(double) * nran, hipMemcpyHostToDevice, ctx->stream));
    }
    double an[3];
    if (next_shift)
        for (int k = 0; k < 3; ++k)
            an[k] = next_shift[k];
    else
        shift_of_step(ctx->p.seed, ctx->step + 1, an);
    if (ctx->md && ctx->md->nb) {  // the swimmer's beads join this collision
        if ((rc = stage_beads_impl(ctx, ctx->md->x[ctx->md->cur], ctx->md->v, ctx->md->nb, hipMemcpyDeviceToDevice,
                                   ctx->md->mobile)))
            return rc;
    }

    const int out = 1 - ctx->cur, par = (int)(ctx->step & 1);
    StepArgs A = base_args(ctx, ctx->cur, tstep);
    A.ovf_cnt = ctx->ovf_cnt + par;
    for (int k = 0; k < 3; ++k) {
        A.a[k] = ctx->a[k];
        A.a_next[k] = an[k];
        A.da[k] = (float)(an[k] - ctx->a[k]);
    }
    A.step = ctx->step;
    A.inj_gamma = inj_gamma ? ctx->inj_gamma : nullptr;
    A.inj_axis = inj_gamma ? ctx->inj_axis : nullptr;
    A.inj_virt = inj_virt ? ctx->inj_virt : nullptr;
    A.inj_ran = inj_ran ? ctx->inj_ran : nullptr;
    A.b_pos3 = ctx->b_pos3;
    A.inj_sph = inj_sph ? ctx->inj_sph : nullptr;
    A.nb = ctx->nb;
    A.b_tile_start = ctx->b_tile_start;
    A.b_lc = ctx->b_lc;
    A.b_idx = ctx->b_idx;
    A.b_vel4 = ctx->b_vel4;
    A.b_out4 = ctx->b_out4;

    launch_begin_step(ctx->diag, false, ctx->stream);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (ctx->timing) {
        HIPCHK(hipEventCreate(&e0));
        HIPCHK(hipEventCreate(&e1));
        HIPCHK(hipEventRecord(e0, ctx->stream));
    }
    // every tile (the timed kernel), then the tiles holding beads, then the
    // tiles above the capacity
    for (int pass : {kPassMain, kPassBeads, kPassLarge}) {
        // FP32-bulk: the tiles a sphere cuts or the wall layers go there too
        if (pass == kPassBeads && ctx->nb == 0 && !(ctx->f32 && (ctx->sphere || ctx->p.walls)))
            continue;
        if (ctx->f32)
            launch_step_f32(A, ctx->variant, ctx->track, pass, ctx->stream);
        else
            launch_step_f64(A, ctx->variant, ctx->track, pass, ctx->stream);
        HIPCHK(hipGetLastError());
        if (ctx->timing && pass == kPassMain) {
            HIPCHK(hipEventRecord(e1, ctx->stream));
            ctx->events.emplace_back(e0, e1);
        }
    }
    launch_fix_overflow(g, ctx->f32, ctx->track, ctx->ovf, ctx->ovf_cnt + par, ctx->ovf_cnt + (1 - par), an,
                        ctx->home[out], ctx->home_cnt[out], ctx->flags, ctx->stream);
    HIPCHK(hipGetLastError());
    if (ctx->sphere) {  // the step's impulses move the sphere
        launch_sphere_finish(ctx->sphere, tstep, g.L, ctx->stream);
        HIPCHK(hipGetLastError());
    }
    if (ctx->md && ctx->md->nb) {  // collided bead velocities back, then the MD half of the step
        HIPCHK(hipMemcpyAsync(ctx->md->v, ctx->b_out4, sizeof(double) * 4 * (size_t)ctx->md->nb,
                              hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHK(md_run(*ctx->md, tstep, ctx->stream));
    }
    ctx->cur = out;
    for (int k = 0; k < 3; ++k)
        ctx->a[k] = an[k];
    ctx->step += 1;
    ctx->nb_last = ctx->nb;
    ctx->nb = 0;
    if (ctx->nb_last)  // beads are consumed by the collision (mpcd.cpp:239-241)
        fill_i32(ctx->b_tile_start, 0, g.ntiles + 1, ctx->stream);
    if (ctx->events.size() >= 256)
        return harvest_events(ctx);
    return MPCD_OK;
}

int mpcd_sphere_attach(mpcd_ctx *ctx, const mpcd_sphere *sp)
{
    if (!ctx || !sp)
        return fail(MPCD_ERR_ARG, "null argument");
    if (int rc = refuse_in_ensemble(ctx))
        return rc;
    if (ctx->p.walls || ctx->p.collision != MPCD_SRD)
        return fail(MPCD_ERR_ARG, "the rigid sphere needs the SRD rule and no walls");
    // the FP32-bulk form: the default tile (its sphere instantiations), and
    // no beads (a cell holding beads and virtual particles would need both
    // to share the residual)
    if (ctx->f32 && ctx->variant != 0)
        return fail(MPCD_ERR_ARG, "the FP32-bulk form takes the rigid sphere with tile variant 0 (the default)");
    if (ctx->f32 && ((ctx->md && ctx->md->nb) || ctx->nb))
        return fail(MPCD_ERR_ARG, "the FP32-bulk form takes the rigid sphere without a swimmer's beads");
    if (ctx->p.rho > kMaxSphereCand)
        return fail(MPCD_ERR_ARG, "rho above the sphere's candidate limit");
    const double half = 0.5 * std::min(ctx->g.L[0], std::min(ctx->g.L[1], ctx->g.L[2]));
    if (!(sp->radius > 0) || sp->radius >= half - 1.0 || !(sp->mass > 0))
        return fail(MPCD_ERR_ARG, "sphere radius must be positive and below half the box minus a cell; mass positive");
    SphereDev h{};
    h.R = sp->radius;
    h.Ms = sp->mass;
    h.Is = sp->inertia > 0 ? sp->inertia : 0.4 * sp->mass * sp->radius * sp->radius;
    for (int k = 0; k < 3; ++k) {
        const double L = ctx->g.L[k];
        h.X[k] = sp->x[k] - L * std::floor(sp->x[k] / L);
        h.V[k] = sp->v[k];
        h.W[k] = sp->w[k];
    }
    h.slip = sp->slip ? 1 : 0;
    int rc;
    if (!ctx->sphere && (rc = dalloc((char **)&ctx->sphere, sizeof(SphereDev))))
        return rc;
    HIPCHK(hipMemcpyAsync(ctx->sphere, &h, sizeof h, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MPCD_OK;
}

int mpcd_sphere_detach(mpcd_ctx *ctx)
{
    if (!ctx)
        return fail(MPCD_ERR_ARG, "null context");
    if (int rc = refuse_in_ensemble(ctx))
        return rc;
    if (ctx->sphere) {
        HIPCHK(hipStreamSynchronize(ctx->stream));
        (void)hipFree(ctx->sphere);
        ctx->sphere = nullptr;
    }
    return MPCD_OK;
}

int mpcd_sphere_state(mpcd_ctx *ctx, mpcd_sphere *out, double force[3], double torque[3])
{
    if (!ctx || !ctx->sphere)
        return fail(MPCD_ERR_STATE, "no sphere attached");
    SphereDev h{};
    HIPCHK(hipMemcpyAsync(&h, ctx->sphere, sizeof h, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (out) {
        out->radius = h.R;
        out->mass = h.Ms;
        out->inertia = h.Is;
        out->slip = h.slip;
        for (int k = 0; k < 3; ++k) {
            out->x[k] = h.X[k];
            out->v[k] = h.V[k];
            out->w[k] = h.W[k];
        }
    }
    for (int k = 0; k < 3; ++k) {
        if (force)
            force[k] = h.h_last > 0 ? h.lastP[k] / h.h_last : 0.0;
        if (torque)
            torque[k] = h.h_last > 0 ? h.lastL[k] / h.h_last : 0.0;
    }
    return MPCD_OK;
}

int mpcd_diagnostics(mpcd_ctx *ctx, mpcd_diag *out)
{
    if (!ctx || !out)
        return fail(MPCD_ERR_ARG, "null argument");
    static thread_local Diag d;
    HIPCHK(hipMemcpyAsync(&d, ctx->diag, sizeof(Diag), hipMemcpyDeviceToHost, ctx->stream));
    int f = 0;
    HIPCHK(hipMemcpyAsync(&f, ctx->flags, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    const double q2 = 1.0 / (2097152.0 * 2097152.0);
    std::memset(out, 0, sizeof *out);
    DiagSlot t{};
    for (int i = 0; i < kDiagSlots; ++i) {
        const DiagSlot &sl = d.slot[i];
        for (int k = 0; k < 3; ++k) {
            t.q_mom[k] += sl.q_mom[k];
            t.mom[k] += sl.mom[k];
        }
        t.q_diss += sl.q_diss;
        t.e2 += sl.e2;
        t.diss += sl.diss;
        t.n += sl.n;
        t.rel_viol += sl.rel_viol;
    }
    if (ctx->f32) {
        for (int k = 0; k < 3; ++k) {
            out->momentum_q[k] = t.q_mom[k];
            out->momentum[k] = (double)t.q_mom[k] / 2097152.0;
        }
    } else {
        for (int k = 0; k < 3; ++k)
            out->momentum[k] = t.mom[k];
    }
    out->twice_kinetic = t.e2;
    out->diss_last = (double)t.q_diss * q2 + t.diss;
    out->diss_total = (double)(d.run_q_diss + t.q_diss) * q2 + d.run_diss + t.diss;
    out->n = ctx->step ? t.n : ctx->n;
    out->rel_violations = d.run_rel_viol + t.rel_viol;
    out->flags = (unsigned)f;
    int load = 0;
    for (int i = 0; i < kDiagSlots; ++i)
        load = std::max(load, d.slot[i].max_load);
    out->deferred_tiles = d.defer_n;
    out->max_tile_load = load;
    out->tile_capacity = (long long)(ctx->f32 ? kF32Variants[ctx->variant].k : kF64Variants[ctx->variant].k) * kBlock;
    return MPCD_OK;
}

int mpcd_reset_diss(mpcd_ctx *ctx)
{
    if (!ctx)
        return fail(MPCD_ERR_ARG, "null context");
    launch_begin_step(ctx->diag, true, ctx->stream);
    HIPCHK(hipGetLastError());
    return MPCD_OK;
}

int mpcd_arm_debug(mpcd_ctx *ctx, int on)
{
    if (!ctx)
        return fail(MPCD_ERR_ARG, "null context");
    if (!ctx->track)
        return fail(MPCD_ERR_STATE, "the debug channel needs track_ids = 1");
    if (int rc = refuse_in_ensemble(ctx))
        return rc;
    if (on) {
        if (!ctx->loaded)
            return fail(MPCD_ERR_STATE, "upload the fluid first");
        if (!ctx->dbg_cell || ctx->dbg_n != ctx->n) {
            (void)hipFree(ctx->dbg_cell);
            (void)hipFree(ctx->dbg_rank);
            int rc;
            if ((rc = dalloc(&ctx->dbg_cell, (size_t)ctx->n)) || (rc = dalloc(&ctx->dbg_rank, (size_t)ctx->n)))
                return rc;
            ctx->dbg_n = ctx->n;
        }
        fill_i32(ctx->dbg_cell, -1, ctx->n, ctx->stream);
        fill_i32(ctx->dbg_rank, -1, ctx->n, ctx->stream);
    } else {
        (void)hipStreamSynchronize(ctx->stream);
        (void)hipFree(ctx->dbg_cell);
        (void)hipFree(ctx->dbg_rank);
        ctx->dbg_cell = ctx->dbg_rank = nullptr;
        ctx->dbg_n = 0;
    }
    return MPCD_OK;
}

int mpcd_debug_binning(mpcd_ctx *ctx, int *cell, int *rank, long long n)
{
    if (!ctx || !ctx->dbg_cell || n != ctx->dbg_n || !cell || !rank)
        return fail(MPCD_ERR_ARG, "debug channel not armed or size mismatch");
    HIPCHK(hipMemcpyAsync(cell, ctx->dbg_cell, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(rank, ctx->dbg_rank, sizeof(int) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MPCD_OK;
}

int mpcd_set_timing(mpcd_ctx *ctx, int on)
{
    if (!ctx)
        return fail(MPCD_ERR_ARG, "null context");
    int rc = harvest_events(ctx);
    ctx->timing = on != 0;
    ctx->timed_ms = 0.0;
    ctx->timed_launches = 0;
    return rc;
}

int mpcd_timing(mpcd_ctx *ctx, double *total_ms, long long *launches)
{
    if (!ctx || !total_ms || !launches)
        return fail(MPCD_ERR_ARG, "null argument");
    int rc = harvest_events(ctx);
    *total_ms = ctx->timed_ms;
    *launches = ctx->timed_launches;
    return rc;
}

int mpcd_tile_variants(int precision, int *shapes4, int max)
{
    const bool f32 = precision == MPCD_FP32BULK;
    if (!f32 && precision != MPCD_FP64)
        return fail(MPCD_ERR_ARG, "unknown precision");
    const int n = f32 ? kNumF32Variants : kNumF64Variants;
    for (int i = 0; i < n && i < max && shapes4; ++i) {
        const TileVariant t = f32 ? kF32Variants[i] : kF64Variants[i];
        shapes4[4 * i + 0] = t.tx;
        shapes4[4 * i + 1] = t.ty;
        shapes4[4 * i + 2] = t.tz;
        shapes4[4 * i + 3] = t.k;
    }
    return n;
}

int mpcd_autotune(const mpcd_params *p, long long n, int steps, int *best_variant, double *best_ms)
{
    if (!p || !best_variant || !best_ms || steps < 1 || n < 0)
        return fail(MPCD_ERR_ARG, "bad autotune arguments");
    const int nvar = mpcd_tile_variants(p->precision, nullptr, 0);
    if (nvar < 0)
        return nvar;
    *best_variant = -1;
    *best_ms = 0.0;
    // the variants that fit, each on a context of its own (warmed up), then
    // three rounds of `steps` timed steps over them in turn, so that a slow
    // moment of the GPU hits them alike; each variant's median round counts.
    // The contexts are held together while device memory allows; a variant
    // that does not fit beside the others is timed in a later group.
    constexpr int kMaxVar = 8, kRounds = 3;
    const int nv = nvar < kMaxVar ? nvar : kMaxVar;
    bool done[kMaxVar] = {};
    int rc = MPCD_OK;
    while (!rc) {
        mpcd_ctx *cs[kMaxVar] = {};
        double ms[kMaxVar][kRounds] = {};
        int held = 0;
        for (int v = 0; v < nv && !rc; ++v) {
            if (done[v])
                continue;
            mpcd_params q = *p;
            q.tile_variant = v;
            q.track_ids = 0;
            if (mpcd_create(&q, &cs[v]) != MPCD_OK) {
                cs[v] = nullptr;
                if (held)
                    continue;  // perhaps memory: a later group
                done[v] = true;  // e.g. rho beyond this variant's capacity
                continue;
            }
            int r2 = mpcd_create_fluid(cs[v], n);
            for (int i = 0; i < 2 && !r2; ++i)
                r2 = mpcd_step(cs[v], 0.1, nullptr, nullptr, nullptr);
            if (r2) {
                mpcd_destroy(cs[v]);
                cs[v] = nullptr;
                if (!held)
                    rc = r2;  // fails alone: an error
                continue;
            }
            ++held;
        }
        if (!held)
            break;
        for (int r = 0; r < kRounds && !rc; ++r) {
            for (int v = 0; v < nv && !rc; ++v) {
                mpcd_ctx *c = cs[v];
                if (!c)
                    continue;
                rc = mpcd_set_timing(c, 1);  // (re)starts the context's timing
                for (int i = 0; i < steps && !rc; ++i)
                    rc = mpcd_step(c, 0.1, nullptr, nullptr, nullptr);
                double t = 0.0;
                long long launches = 0;
                if (!rc)
                    rc = mpcd_timing(c, &t, &launches);
                if (!rc)
                    rc = check_flags(c);
                ms[v][r] = t / (double)(launches ? launches : 1);
            }
        }
        for (int v = 0; v < nv; ++v) {
            if (!cs[v])
                continue;
            mpcd_destroy(cs[v]);
            done[v] = true;
            if (rc)
                continue;
            double *m = ms[v];
            std::sort(m, m + kRounds);
            if (*best_variant < 0 || m[kRounds / 2] < *best_ms) {
                *best_variant = v;
                *best_ms = m[kRounds / 2];
            }
        }
    }
    if (rc)
        return rc;
    if (*best_variant < 0)
        return fail(MPCD_ERR_ARG, "no tile variant fits these parameters");
    return MPCD_OK;
}

int mpcd_sync(mpcd_ctx *ctx)
{
    if (!ctx)
        return fail(MPCD_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return check_flags(ctx);
}


/* ------------------------------------------------------------ bead MD */
namespace {

// Per-bead ordered force terms from the object description: each bead's
// list holds its contributions in the order the reference's loops add
// them -- Sphere::recalcforce (sphere.cpp:87-119), Cilia::recalcforce and
// addcross (cilia.cpp:127-239), then Clami::global_forces (clami.cpp:188-198).
int md_build(const mpcd_md_desc *d, std::vector<int> &tstart, std::vector<MdTerm> &terms,
             std::vector<double> &ktab, std::vector<MdCilium> &cil, std::vector<double> &nl, std::vector<double> &ol)
{
    const int nb = d->nb;
    std::vector<std::vector<MdTerm>> per((size_t)nb);
    bool ok = true, kfull = false;
    // c: the cilium of a rest-length term (the list order is the record)
    auto add = [&](int i, int j, int flags, int c, int idx, double b, double k) {
        (void)c;
        if (i < 0 || i >= nb || j < 0 || j >= nb || idx < 0 || idx >= kTermMaxIdx) {
            ok = false;
            return;
        }
        // the spring constant's index in the table of distinct constants
        // (equal bits: the device reads back the same double)
        size_t ki = 0;
        while (ki < ktab.size() && std::memcmp(&ktab[ki], &k, sizeof k) != 0)
            ++ki;
        if (ki == ktab.size()) {
            if (ki == (size_t)kTermMaxK) {
                kfull = true;
                return;
            }
            ktab.push_back(k);
        }
        const uint32_t meta = (uint32_t)flags | ((uint32_t)ki << kTermKShift) | ((uint32_t)idx << kTermIdxShift);
        per[(size_t)i].push_back(MdTerm{j, meta, b});
    };
    for (int s = 0; s < d->nsph; ++s) {
        const mpcd_md_sphere &S = d->sph[s];
        const int o = S.first;
        const double kr = 0.1 * S.springconstant;
        for (int i = 1; i < S.npart; i++) {
            add(o, o + i, kTermConst, 0, 0, S.radius, kr);
            add(o + i, o, kTermConst, 0, 0, S.radius, kr);
        }
        for (int i = 1; i < S.npart; i++)
            for (int n = d->nei_start[S.nei_off + i]; n < d->nei_start[S.nei_off + i + 1]; ++n)
                add(o + i, o + d->nei[n], kTermConst, 0, 0, S.headbond, S.springconstant);
    }
    int off = 0;
    for (int c = 0; c < d->ncil; ++c) {
        const mpcd_md_cilium &C = d->cil[c];
        const int R = C.rodlength, o = C.first;
        if (R < 2 || C.num_anker < 1 || o < 0 || o + 3 * R > nb)
            return fail(MPCD_ERR_ARG, "cilium outside the bead array");
        const double bl = C.bondlength, k = C.springconstant, ck = 1.0 * k, cross = std::sqrt(2.0) * bl;
        for (int i = 0; i < R - 1; ++i) {
            add(o + i, o + i + 1, kTermNewLen, c, off + i, bl, k);
            add(o + i + 1, o + i, kTermNewLen, c, off + i, bl, k);
            add(o + R + i, o + R + i + 1, kTermConst, 0, 0, bl, k);
            add(o + R + i + 1, o + R + i, kTermConst, 0, 0, bl, k);
            add(o + 2 * R + i, o + 2 * R + i + 1, kTermConst, 0, 0, bl, k);
            add(o + 2 * R + i + 1, o + 2 * R + i, kTermConst, 0, 0, bl, k);
        }
        for (int i = 0; i < R; ++i) {
            add(o + i, o + i + R, kTermConst, 0, 0, bl, k);
            add(o + i, o + i + 2 * R, kTermConst, 0, 0, bl, k);
            add(o + i + R, o + i, kTermConst, 0, 0, bl, k);
            add(o + i + R, o + i + 2 * R, kTermConst, 0, 0, bl, k);
            add(o + i + 2 * R, o + i, kTermConst, 0, 0, bl, k);
            add(o + i + 2 * R, o + i + R, kTermConst, 0, 0, bl, k);
        }
        for (int i = 0; i < R - 1; ++i) {
            add(o + i, o + i + R + 1, kTermDiag, c, off + i, bl, ck);
            add(o + i, o + i + 2 * R + 1, kTermDiag, c, off + i, bl, ck);
            add(o + i + R, o + i + 1, kTermDiag, c, off + i, bl, ck);
            add(o + i + 2 * R, o + i + 1, kTermDiag, c, off + i, bl, ck);
            add(o + i + R, o + i + 2 * R + 1, kTermConst, 0, 0, cross, ck);
            add(o + i + 2 * R, o + i + R + 1, kTermConst, 0, 0, cross, ck);
        }
        for (int i = 1; i < R; ++i) {
            add(o + i, o + i + R - 1, kTermDiag, c, off + i - 1, bl, ck);
            add(o + i, o + i + 2 * R - 1, kTermDiag, c, off + i - 1, bl, ck);
            add(o + i + R, o + i - 1, kTermDiag, c, off + i - 1, bl, ck);
            add(o + i + 2 * R, o + i - 1, kTermDiag, c, off + i - 1, bl, ck);
            add(o + i + R, o + i + 2 * R - 1, kTermConst, 0, 0, cross, ck);
            add(o + i + 2 * R, o + i + R - 1, kTermConst, 0, 0, cross, ck);
        }
        MdCilium m{};
        m.first = o;
        m.rodlength = R;
        m.num_anker = C.num_anker;
        m.constant_beat = C.constant_beat ? 1 : 0;
        m.len_off = off;
        m.bondlength = bl;
        m.beatstrength = C.beatstrength;
        m.bending = C.bending;
        m.min_bend = C.min_bend;
        m.max_bend = C.max_bend;
        m.m_speed = C.m_speed;
        m.springconstant = k;
        m.power_stroke_duration = C.power_stroke_duration;
        m.recovery_stroke_duration = C.recovery_stroke_duration;
        m.midpnt = C.midpnt;
        m.Tstroke = C.Tstroke;
        cil.push_back(m);
        for (int i = 0; i < R; ++i) {
            nl.push_back(d->newlength[off + i]);
            ol.push_back(d->oldlength[off + i]);
        }
        off += R;
    }
    for (int a = 0; a < d->natt; ++a) {
        const mpcd_md_attachment &T = d->att[a];
        add(T.bead0, T.bead1, kTermConst | kTermAttach, 0, 0, T.l0, d->att_springconstant);
        add(T.bead1, T.bead0, kTermConst | kTermAttach | kTermAttachB, 0, 0, T.l0, d->att_springconstant);
    }
    if (kfull)
        return fail(MPCD_ERR_ARG, "more than 256 distinct spring constants in one swimmer");
    if (!ok)
        return fail(MPCD_ERR_ARG, "spring bead index outside the bead array, or too many rest lengths");
    tstart.assign((size_t)nb + 1, 0);
    for (int b = 0; b < nb; ++b) {
        tstart[(size_t)b + 1] = tstart[(size_t)b] + (int)per[(size_t)b].size();
        terms.insert(terms.end(), per[(size_t)b].begin(), per[(size_t)b].end());
    }
    return MPCD_OK;
}

}  // namespace

int mpcd_md_attach(mpcd_ctx *ctx, const mpcd_md_desc *d)
{
    if (!ctx || !d || d->nb < 0 || (d->nb && (!d->pos3 || !d->vel4)) || d->nsph < 0 || d->ncil < 0 || d->natt < 0 ||
        (d->nsph && (!d->sph || !d->nei_start)) || (d->ncil && (!d->cil || !d->newlength || !d->oldlength)) ||
        (d->natt && !d->att) || !(d->ststep > 0))
        return fail(MPCD_ERR_ARG, "bad swimmer description");
    if (int rc = refuse_in_ensemble(ctx))
        return rc;
    std::vector<int> tstart;
    std::vector<MdTerm> terms;
    std::vector<double> ktab;
    std::vector<MdCilium> cil;
    std::vector<double> nl, ol;
    int rc = md_build(d, tstart, terms, ktab, cil, nl, ol);
    if (rc)
        return rc;
    if (ctx->md) {
        md_free(*ctx->md);
    } else {
        ctx->md = new MdSystem;
    }
    MdSystem &s = *ctx->md;
    s.nb = d->nb;
    s.nterms = (int)terms.size();
    s.ncil = (int)cil.size();
    s.nlen = (int)nl.size();
    s.globalt = d->globalt;
    s.ststep = d->ststep;
    s.walls = ctx->p.walls;
    s.ztop = ctx->p.size[2] - 1.0;
    s.vw[0] = ctx->p.wall_v[0];
    s.vw[1] = ctx->p.wall_v[1];
    if (md_alloc_upload(s, d->pos3, d->vel4, d->mobile, tstart, terms, ktab, cil, nl, ol, ctx->stream) != hipSuccess) {
        md_free(s);
        delete ctx->md;
        ctx->md = nullptr;
        return fail(MPCD_ERR_HIP, "device allocation for the swimmer failed");
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return MPCD_OK;
}

int mpcd_md_detach(mpcd_ctx *ctx)
{
    if (!ctx)
        return fail(MPCD_ERR_ARG, "null context");
    if (int rc = refuse_in_ensemble(ctx))
        return rc;
    if (ctx->md) {
        (void)hipStreamSynchronize(ctx->stream);
        md_free(*ctx->md);
        delete ctx->md;
        ctx->md = nullptr;
    }
    return MPCD_OK;
}

int mpcd_md_run(mpcd_ctx *ctx, double tstep)
{
    if (!ctx || !ctx->md)
        return fail(MPCD_ERR_STATE, "no swimmer attached");
    if (!(tstep > 0))
        return fail(MPCD_ERR_ARG, "tstep must be positive");
    if (int rc = refuse_in_ensemble(ctx))
        return rc;
    HIPCHK(md_run(*ctx->md, tstep, ctx->stream));
    return MPCD_OK;
}

int mpcd_md_fetch(mpcd_ctx *ctx, double *pos3, double *vel4, double *force3, double *beat2, double *newlength,
                  double *oldlength, double *globalt)
{
    if (!ctx || !ctx->md)
        return fail(MPCD_ERR_STATE, "no swimmer attached");
    MdSystem &s = *ctx->md;
    const size_t nb = (size_t)s.nb;
    if (pos3 && nb)
        HIPCHK(hipMemcpyAsync(pos3, s.x[s.cur], sizeof(double) * 3 * nb, hipMemcpyDeviceToHost, ctx->stream));
    if (vel4 && nb)
        HIPCHK(hipMemcpyAsync(vel4, s.v, sizeof(double) * 4 * nb, hipMemcpyDeviceToHost, ctx->stream));
    if (force3 && nb)
        HIPCHK(hipMemcpyAsync(force3, s.f, sizeof(double) * 3 * nb, hipMemcpyDeviceToHost, ctx->stream));
    std::vector<MdCilium> cil((size_t)s.ncil);
    if (s.ncil)
        HIPCHK(hipMemcpyAsync(cil.data(), s.cil, sizeof(MdCilium) * cil.size(), hipMemcpyDeviceToHost, ctx->stream));
    if (newlength && s.nlen)
        HIPCHK(hipMemcpyAsync(newlength, s.newlen, sizeof(double) * (size_t)s.nlen, hipMemcpyDeviceToHost, ctx->stream));
    if (oldlength && s.nlen)
        HIPCHK(hipMemcpyAsync(oldlength, s.oldlen, sizeof(double) * (size_t)s.nlen, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (beat2)
        for (size_t c = 0; c < cil.size(); ++c) {
            beat2[2 * c] = cil[c].midpnt;
            beat2[2 * c + 1] = cil[c].Tstroke;
        }
    if (globalt)
        *globalt = s.globalt;
    return MPCD_OK;
}

// ---------------------------------------------------------------- checkpoints
}  // extern "C"

namespace {

constexpr char kCkptMagic[8] = {'M', 'C', 'M', 'C', 'K', 'P', 'T', '1'};

struct CkptHeader {
    char magic[8];
    int32_t size[3], rho, precision, track, walls, collision, words;
    uint64_t seed, step;
    double a[3];
    int64_t n;
    int64_t run_q_diss;
    double run_diss;
    int64_t run_rel_viol;
    int32_t has_sphere, has_md;
    int32_t md_nb, md_ncil, md_nlen, pad;
    double md_globalt;
};

struct FileCloser {
    FILE *f;
    ~FileCloser()
    {
        if (f)
            fclose(f);
    }
};

template <class T>
bool put(FILE *f, const T *p, size_t count)
{
    return count == 0 || fwrite(p, sizeof(T), count, f) == count;
}

template <class T>
bool get(FILE *f, T *p, size_t count)
{
    return count == 0 || fread(p, sizeof(T), count, f) == count;
}

// The records in tile order, as stored (W words each).
int download_raw(mpcd_ctx *c, std::vector<uint32_t> &words)
{
    const Geom &g = c->g;
    int *tot = nullptr;
    long long *prefix = nullptr;
    uint32_t *dw = nullptr;
    int rc;
    if ((rc = dalloc(&tot, (size_t)g.ntiles)) || (rc = dalloc(&prefix, (size_t)g.ntiles)) ||
        (rc = dalloc(&dw, (size_t)c->n * c->W))) {
        (void)hipFree(tot);
        (void)hipFree(prefix);
        return rc;
    }
    std::vector<int> htot((size_t)g.ntiles);
    std::vector<long long> hpre((size_t)g.ntiles);
    launch_tile_totals(g, c->home_cnt[c->cur], c->ob_cnt[c->cur], tot, c->stream);
    rc = MPCD_OK;
    if (hipMemcpyAsync(htot.data(), tot, sizeof(int) * g.ntiles, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
        rc = fail(MPCD_ERR_HIP, "checkpoint totals failed");
    long long run = 0;
    for (int t = 0; t < g.ntiles && !rc; ++t) {
        hpre[(size_t)t] = run;
        run += htot[(size_t)t];
    }
    if (!rc && run != c->n)
        rc = fail(MPCD_ERR_STATE, "device holds " + std::to_string(run) + " particles, expected " + std::to_string(c->n));
    if (!rc && hipMemcpyAsync(prefix, hpre.data(), sizeof(long long) * g.ntiles, hipMemcpyHostToDevice, c->stream) !=
                   hipSuccess)
        rc = fail(MPCD_ERR_HIP, "checkpoint prefix copy failed");
    if (!rc) {
        launch_download_raw(g, c->f32, c->track, c->home[c->cur], c->ob[c->cur], c->home_cnt[c->cur],
                            c->ob_cnt[c->cur], prefix, dw, c->stream);
        words.resize((size_t)c->n * c->W);
        if ((c->n && hipMemcpyAsync(words.data(), dw, sizeof(uint32_t) * words.size(), hipMemcpyDeviceToHost,
                                    c->stream) != hipSuccess) ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = fail(MPCD_ERR_HIP, "checkpoint copy failed");
    }
    (void)hipFree(tot);
    (void)hipFree(prefix);
    (void)hipFree(dw);
    return rc;
}

}  // namespace

extern "C" {

int mpcd_checkpoint_save(mpcd_ctx *ctx, const char *path)
{
    if (!ctx || !path)
        return fail(MPCD_ERR_ARG, "null argument");
    if (!ctx->loaded)
        return fail(MPCD_ERR_STATE, "no fluid uploaded");
    if (ctx->nb)
        return fail(MPCD_ERR_STATE, "beads are staged for the next collision: checkpoint between steps");
    int rc = check_flags(ctx);
    if (rc)
        return rc;
    CkptHeader h{};
    std::memcpy(h.magic, kCkptMagic, 8);
    for (int k = 0; k < 3; ++k) {
        h.size[k] = ctx->p.size[k];
        h.a[k] = ctx->a[k];
    }
    h.rho = ctx->p.rho;
    h.precision = ctx->p.precision;
    h.track = ctx->track;
    h.walls = ctx->p.walls;
    h.collision = ctx->p.collision;
    h.words = ctx->W;
    h.seed = ctx->p.seed;
    h.step = ctx->step;
    h.n = ctx->n;
    Diag d;
    HIPCHK(hipMemcpyAsync(&d, ctx->diag, sizeof(Diag), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // the running totals as the next step's fold will leave them
    long long q = d.run_q_diss, rv = d.run_rel_viol;
    double dd = d.run_diss;
    for (int i = 0; i < kDiagSlots; ++i) {
        q += d.slot[i].q_diss;
        dd += d.slot[i].diss;
        rv += d.slot[i].rel_viol;
    }
    h.run_q_diss = q;
    h.run_diss = dd;
    h.run_rel_viol = rv;
    h.has_sphere = ctx->sphere != nullptr;
    MdSystem *md = (ctx->md && ctx->md->nb) ? ctx->md : nullptr;
    h.has_md = md != nullptr;
    if (md) {
        h.md_nb = md->nb;
        h.md_ncil = md->ncil;
        h.md_nlen = md->nlen;
        h.md_globalt = md->globalt;
    }
    std::vector<uint32_t> words;
    if ((rc = download_raw(ctx, words)))
        return rc;
    SphereDev sp{};
    if (ctx->sphere)
        HIPCHK(hipMemcpy(&sp, ctx->sphere, sizeof sp, hipMemcpyDeviceToHost));
    std::vector<double> x, v, f, nl, ol;
    std::vector<MdCilium> cil;
    if (md) {
        x.resize((size_t)md->nb * 3);
        v.resize((size_t)md->nb * 4);
        f.resize((size_t)md->nb * 3);
        cil.resize((size_t)md->ncil);
        nl.resize((size_t)md->nlen);
        ol.resize((size_t)md->nlen);
        HIPCHK(hipMemcpy(x.data(), md->x[md->cur], sizeof(double) * x.size(), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(v.data(), md->v, sizeof(double) * v.size(), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(f.data(), md->f, sizeof(double) * f.size(), hipMemcpyDeviceToHost));
        if (md->ncil)
            HIPCHK(hipMemcpy(cil.data(), md->cil, sizeof(MdCilium) * cil.size(), hipMemcpyDeviceToHost));
        if (md->nlen) {
            HIPCHK(hipMemcpy(nl.data(), md->newlen, sizeof(double) * nl.size(), hipMemcpyDeviceToHost));
            HIPCHK(hipMemcpy(ol.data(), md->oldlen, sizeof(double) * ol.size(), hipMemcpyDeviceToHost));
        }
    }
    FileCloser fc{fopen(path, "wb")};
    if (!fc.f)
        return fail(MPCD_ERR_ARG, std::string("cannot open for writing: ") + path);
    bool ok = put(fc.f, &h, 1) && put(fc.f, words.data(), words.size());
    if (ok && ctx->sphere)
        ok = put(fc.f, &sp, 1);
    if (ok && md)
        ok = put(fc.f, x.data(), x.size()) && put(fc.f, v.data(), v.size()) && put(fc.f, f.data(), f.size()) &&
             put(fc.f, cil.data(), cil.size()) && put(fc.f, nl.data(), nl.size()) && put(fc.f, ol.data(), ol.size());
    if (!ok || fflush(fc.f) != 0)
        return fail(MPCD_ERR_HIP, std::string("write failed: ") + path);
    return MPCD_OK;
}

int mpcd_checkpoint_load(mpcd_ctx *ctx, const char *path)
{
    if (!ctx || !path)
        return fail(MPCD_ERR_ARG, "null argument");
    if (int rc = refuse_in_ensemble(ctx))
        return rc;
    FileCloser fc{fopen(path, "rb")};
    if (!fc.f)
        return fail(MPCD_ERR_ARG, std::string("cannot open: ") + path);
    CkptHeader h{};
    if (!get(fc.f, &h, 1) || std::memcmp(h.magic, kCkptMagic, 8) != 0)
        return fail(MPCD_ERR_ARG, "not a checkpoint of this library");
    for (int k = 0; k < 3; ++k)
        if (h.size[k] != ctx->p.size[k])
            return fail(MPCD_ERR_ARG, "checkpoint box size differs from the context's");
    if (h.rho != ctx->p.rho || h.precision != ctx->p.precision || h.track != (int)ctx->track ||
        h.walls != ctx->p.walls || h.collision != ctx->p.collision || h.words != ctx->W)
        return fail(MPCD_ERR_ARG, "checkpoint parameters (rho, precision, track_ids, walls, collision) differ");
    if ((h.has_sphere != 0) != (ctx->sphere != nullptr))
        return fail(MPCD_ERR_STATE, "attach the sphere (any state) before loading a checkpoint that has one, and only then");
    MdSystem *md = (ctx->md && ctx->md->nb) ? ctx->md : nullptr;
    if ((h.has_md != 0) != (md != nullptr))
        return fail(MPCD_ERR_STATE, "attach the same swimmer before loading a checkpoint that has one, and only then");
    if (md && (h.md_nb != md->nb || h.md_ncil != md->ncil || h.md_nlen != md->nlen))
        return fail(MPCD_ERR_ARG, "the attached swimmer's sizes differ from the checkpoint's");
    if (h.n < 0 || (ctx->track && h.n > 0x7fffffffLL))
        return fail(MPCD_ERR_ARG, "bad particle count");
    std::vector<uint32_t> words((size_t)h.n * h.words);
    if (!get(fc.f, words.data(), words.size()))
        return fail(MPCD_ERR_ARG, "truncated checkpoint (records)");
    SphereDev sp{};
    if (h.has_sphere && !get(fc.f, &sp, 1))
        return fail(MPCD_ERR_ARG, "truncated checkpoint (sphere)");
    std::vector<double> x, v, f, nl, ol;
    std::vector<MdCilium> cil;
    if (md) {
        x.resize((size_t)md->nb * 3);
        v.resize((size_t)md->nb * 4);
        f.resize((size_t)md->nb * 3);
        cil.resize((size_t)md->ncil);
        nl.resize((size_t)md->nlen);
        ol.resize((size_t)md->nlen);
        if (!(get(fc.f, x.data(), x.size()) && get(fc.f, v.data(), v.size()) && get(fc.f, f.data(), f.size()) &&
              get(fc.f, cil.data(), cil.size()) && get(fc.f, nl.data(), nl.size()) && get(fc.f, ol.data(), ol.size())))
            return fail(MPCD_ERR_ARG, "truncated checkpoint (swimmer)");
    }
    // the solvent, binned under the saved shift
    const Geom &g = ctx->g;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    uint32_t *dw = nullptr;
    int rc;
    if ((rc = dalloc(&dw, words.size())))
        return rc;
    for (int k = 0; k < 3; ++k)
        ctx->a[k] = h.a[k];
    ctx->p.seed = h.seed;
    ctx->step = h.step;
    ctx->nb = 0;
    fill_i32(ctx->home_cnt[ctx->cur], 0, g.ntiles, ctx->stream);
    fill_i32(ctx->ob_cnt[ctx->cur], 0, (long long)g.ntiles * 26, ctx->stream);
    fill_i32(ctx->ovf_cnt, 0, 2, ctx->stream);
    fill_i32(ctx->b_tile_start, 0, g.ntiles + 1, ctx->stream);
    if (h.n && hipMemcpyAsync(dw, words.data(), sizeof(uint32_t) * words.size(), hipMemcpyHostToDevice, ctx->stream) !=
                   hipSuccess) {
        (void)hipFree(dw);
        return fail(MPCD_ERR_HIP, "checkpoint upload failed");
    }
    launch_upload_raw(g, ctx->f32, ctx->track, dw, h.n, ctx->a, ctx->home[ctx->cur], ctx->home_cnt[ctx->cur],
                      ctx->flags, ctx->stream);
    const hipError_t le = hipGetLastError();
    const hipError_t se = hipStreamSynchronize(ctx->stream);
    (void)hipFree(dw);
    if (le != hipSuccess || se != hipSuccess)
        return fail(MPCD_ERR_HIP, "checkpoint upload failed");
    ctx->n = h.n;
    ctx->loaded = true;
    // the running totals, per-step slots cleared
    fill_u64(ctx->diag, 0ull, (long long)(sizeof(Diag) / 8), ctx->stream);
    Diag *dd = ctx->diag;
    HIPCHK(hipMemcpyAsync(&dd->run_q_diss, &h.run_q_diss, sizeof(long long), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(&dd->run_diss, &h.run_diss, sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(&dd->run_rel_viol, &h.run_rel_viol, sizeof(long long), hipMemcpyHostToDevice, ctx->stream));
    if (h.has_sphere)
        HIPCHK(hipMemcpyAsync(ctx->sphere, &sp, sizeof sp, hipMemcpyHostToDevice, ctx->stream));
    if (md) {
        HIPCHK(hipMemcpyAsync(md->x[md->cur], x.data(), sizeof(double) * x.size(), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(md->v, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(md->f, f.data(), sizeof(double) * f.size(), hipMemcpyHostToDevice, ctx->stream));
        if (md->ncil)
            HIPCHK(hipMemcpyAsync(md->cil, cil.data(), sizeof(MdCilium) * cil.size(), hipMemcpyHostToDevice,
                                  ctx->stream));
        if (md->nlen) {
            HIPCHK(hipMemcpyAsync(md->newlen, nl.data(), sizeof(double) * nl.size(), hipMemcpyHostToDevice, ctx->stream));
            HIPCHK(hipMemcpyAsync(md->oldlen, ol.data(), sizeof(double) * ol.size(), hipMemcpyHostToDevice, ctx->stream));
        }
        md->globalt = h.md_globalt;
        md->graph_tstep = -1.0;  // its device state restarts from this one
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return check_flags(ctx);
}

// ------------------------------------------------------------------ ensembles
namespace {

// The launches of one ensemble step, the wide MD path's refreshes at the
// evaluations of sched; every argument is a device table, so the same
// launches replay from a graph.
int ens_enqueue(mpcd_ensemble *e, const MdSchedule &sched)
{
    const int nbox = (int)e->box.size();
    launch_ens_begin(e->d_boxes, e->d_args, nbox, e->stream);
    // FP32-bulk: the tiles a sphere cuts go to the bead pass too
    for (int pass : {kPassMain, kPassBeads, kPassLarge})
        if (pass != kPassBeads || e->beads || (e->f32 && e->sphere))
            launch_step_multi(e->f32, e->d_args, e->d_tile_start, nbox, e->ntiles, e->variant, e->track, pass,
                              e->stream, e->rule, e->sphere);
    launch_ens_fix(e->d_boxes, e->d_args, nbox, e->f32, e->track, e->stream);
    if (e->md_wide)
        md_launch_multi_wide(e->d_md, e->wide, e->tstep, e->stream, &sched);
    else
        md_launch_multi(e->d_md, e->nsw, e->tstep, e->stream);
    HIPCHK(hipGetLastError());
    return MPCD_OK;
}

// Device tables for steps of length tstep, from the members' host state
// (the device is idle: the caller synchronised); the graphs captured for
// another state or length go.
int ens_build(mpcd_ensemble *e, double tstep)
{
    e->graphs.clear();
    const int nbox = (int)e->box.size();
    std::vector<EnsBox> hb((size_t)nbox);
    std::vector<MdBox> hm;
    std::vector<int> ts((size_t)nbox + 1, 0);
    for (int b = 0; b < nbox; ++b) {
        mpcd_ctx *c = e->box[(size_t)b];
        if (c->md && c->md->nb) {
            const int n = md_substeps(c->md->ststep, tstep);
            if (n < 1 || n >= kMdMaxSubsteps)
                return fail(MPCD_ERR_ARG, "tstep / ststep out of the MD sub-step range");
        }
    }
    for (int b = 0; b < nbox; ++b) {
        mpcd_ctx *c = e->box[(size_t)b];
        const bool sw = c->md && c->md->nb;
        if (sw)
            hm.push_back(md_box(*c->md, tstep));
        EnsBox &B = hb[(size_t)b];
        for (int cur = 0; cur < 2; ++cur) {
            StepArgs A = base_args(c, cur, tstep);
            A.b_tile_start = c->b_tile_start;
            if (sw) {
                A.nb = c->md->nb;
                A.b_lc = c->b_lc;
                A.b_idx = c->b_idx;
                A.b_vel4 = c->b_vel4;
                A.b_out4 = c->md->v;  // collided bead velocities straight into the swimmer
            }
            B.tmpl[cur] = A;
        }
        B.ovf_cnt2 = c->ovf_cnt;
        B.step = c->step;
        for (int k = 0; k < 3; ++k)
            B.a[k] = c->a[k];
        B.cur = c->cur;
        B.md = sw ? e->d_md + (hm.size() - 1) : nullptr;
        ts[(size_t)b + 1] = ts[(size_t)b] + c->g.ntiles;
    }
    e->nsw = (int)hm.size();
    e->ntiles = ts[(size_t)nbox];
    // the wide MD path: each phase's workgroups split among the swimmers
    {
        std::vector<int> tab;
        md_wide_tables(hm.data(), e->nsw, tab, e->wide, e->md_block);
        (void)hipFree(e->d_mdblk);
        e->d_mdblk = nullptr;
        int rc;
        if ((rc = dalloc(&e->d_mdblk, tab.size())))
            return rc;
        HIPCHK(hipMemcpy(e->d_mdblk, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice));
        md_wide_bind(e->wide, e->d_mdblk);
    }
    HIPCHK(hipMemcpy(e->d_boxes, hb.data(), sizeof(EnsBox) * hb.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->d_tile_start, ts.data(), sizeof(int) * ts.size(), hipMemcpyHostToDevice));
    if (!hm.empty())
        HIPCHK(hipMemcpy(e->d_md, hm.data(), sizeof(MdBox) * hm.size(), hipMemcpyHostToDevice));
    e->tstep = tstep;
    return MPCD_OK;
}

// The graph of one ensemble step with the wide MD path's refreshes at sched.
hipError_t ens_capture(mpcd_ensemble *e, const MdSchedule &sched, hipGraph_t &g, hipGraphExec_t &x)
{
    hipError_t ec = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal);
    if (ec != hipSuccess)
        return ec;
    const int rc = ens_enqueue(e, sched);
    ec = hipStreamEndCapture(e->stream, &g);
    if (rc || ec != hipSuccess) {
        if (g)
            (void)hipGraphDestroy(g);
        g = nullptr;
        return ec != hipSuccess ? ec : hipErrorLaunchFailure;
    }
    if ((ec = hipGraphInstantiate(&x, g, nullptr, nullptr, 0)) != hipSuccess) {
        (void)hipGraphDestroy(g);
        g = nullptr;
        x = nullptr;
    }
    return ec;
}

// The union of the swimmers' refresh schedules for the next step of length
// tstep, from the host's view of their times (the device advances them by
// the same additions, so each swimmer's own schedule is inside the union).
MdSchedule ens_schedule(const mpcd_ensemble *e, double tstep)
{
    MdSchedule u{};
    for (const mpcd_ctx *c : e->box) {
        if (!(c->md && c->md->nb))
            continue;
        MdSchedule s;
        md_schedule(c->md->globalt, c->md->ststep, md_substeps(c->md->ststep, tstep), c->md->ncil > 0, s);
        for (int i = 0; i < kMdMaxSubsteps / 32; ++i)
            u.w[i] |= s.w[i];
    }
    return u;
}

// What one step does to a member's host state (the device did the step).
void ens_advance_host(mpcd_ctx *c, double tstep)
{
    c->cur = 1 - c->cur;
    c->step += 1;
    shift_of_step(c->p.seed, c->step, c->a);
    c->nb = 0;
    c->nb_last = 0;
    if (c->md && c->md->nb) {
        md_advance_host(*c->md, tstep);
        c->md->graph_tstep = -1.0;  // its own MD graph restarts from this state
    }
}

}  // namespace

int mpcd_ensemble_create(mpcd_ctx *const *boxes, int nbox, mpcd_ensemble **out)
{
    if (!boxes || !out || nbox < 1)
        return fail(MPCD_ERR_ARG, "null argument or no boxes");
    *out = nullptr;
    for (int b = 0; b < nbox; ++b) {
        mpcd_ctx *c = boxes[b];
        if (!c)
            return fail(MPCD_ERR_ARG, "null context in the ensemble");
        for (int q = 0; q < b; ++q)
            if (boxes[q] == c)
                return fail(MPCD_ERR_ARG, "a context appears twice in the ensemble");
        if (c->ens)
            return fail(MPCD_ERR_STATE, "a context is already in an ensemble");
        if (c->f32 != boxes[0]->f32 || c->track != boxes[0]->track || c->variant != boxes[0]->variant ||
            c->p.collision != boxes[0]->p.collision)
            return fail(MPCD_ERR_ARG, "ensemble members must share precision, track_ids, tile variant and collision rule");
        if ((c->sphere != nullptr) != (boxes[0]->sphere != nullptr))
            return fail(MPCD_ERR_ARG, "ensemble members all have a rigid sphere, or none has");
        if (c->f32 && (c->p.walls || (c->sphere && c->md && c->md->nb)))
            return fail(MPCD_ERR_ARG, "FP32-bulk ensemble members take no walls, nor a swimmer beside a rigid sphere");
        if (c->p.collision != MPCD_SRD && c->variant != kAtVariant)
            return fail(MPCD_ERR_ARG, "MPC-AT ensemble members need the MPC-AT tile variant");
        if (!c->loaded)
            return fail(MPCD_ERR_STATE, "every member needs its fluid loaded");
        if (c->nb)
            return fail(MPCD_ERR_STATE, "a member has staged beads");
        if (c->md && c->md->nb > kEnsMaxBeads)
            return fail(MPCD_ERR_ARG, "a swimmer of an ensemble has at most 16384 beads");
    }
    mpcd_ensemble *e = new mpcd_ensemble();
    auto bail = [&](int rc) {
        (void)hipFree(e->d_boxes);
        (void)hipFree(e->d_args);
        (void)hipFree(e->d_tile_start);
        (void)hipFree(e->d_md);
        if (e->stream)
            (void)hipStreamDestroy(e->stream);
        delete e;
        return rc;
    };
    e->f32 = boxes[0]->f32;
    e->track = boxes[0]->track;
    e->variant = boxes[0]->variant;
    e->rule = boxes[0]->p.collision == MPCD_AT ? kRuleAt : (boxes[0]->p.collision == MPCD_AT_A ? kRuleAtA : kRuleSrd);
    e->sphere = boxes[0]->sphere != nullptr;
    int nsw = 0;
    for (int b = 0; b < nbox; ++b)
        if (boxes[b]->md && boxes[b]->md->nb)
            ++nsw;
    e->beads = nsw > 0;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        e->stream = nullptr;
        return bail(fail(MPCD_ERR_HIP, "hipStreamCreate failed"));
    }
    int rc;
    if ((rc = dalloc(&e->d_boxes, (size_t)nbox)) || (rc = dalloc(&e->d_args, (size_t)nbox)) ||
        (rc = dalloc(&e->d_tile_start, (size_t)nbox + 1)) || (rc = dalloc(&e->d_md, (size_t)nsw)))
        return bail(rc);
    for (int b = 0; b < nbox; ++b) {
        mpcd_ctx *c = boxes[b];
        if (c->md && c->md->nb && (rc = ensure_bead_cap(c, c->md->nb)))
            return bail(rc);
    }
    for (int b = 0; b < nbox; ++b) {
        mpcd_ctx *c = boxes[b];
        if (hipStreamSynchronize(c->stream) != hipSuccess)
            return bail(fail(MPCD_ERR_HIP, "a member's stream failed"));
        fill_i32(c->b_tile_start, 0, c->g.ntiles + 1, e->stream);
    }
    for (int b = 0; b < nbox; ++b) {
        mpcd_ctx *c = boxes[b];
        e->box.push_back(c);
        e->own_stream.push_back(c->stream);
        c->stream = e->stream;
        c->ens = e;
    }
    *out = e;
    return MPCD_OK;
}

int mpcd_ensemble_set_md_wide(mpcd_ensemble *e, int on)
{
    if (!e)
        return fail(MPCD_ERR_ARG, "null ensemble");
    if (e->md_wide != (on != 0)) {
        e->md_wide = on != 0;
        e->tstep = -1.0;  // rebuilt at the next step
    }
    return MPCD_OK;
}

int mpcd_ensemble_set_md_block(mpcd_ensemble *e, int threads)
{
    if (!e)
        return fail(MPCD_ERR_ARG, "null ensemble");
    if (!md_block_ok(threads))
        return fail(MPCD_ERR_ARG, "MD workgroup size must be 256, 512 or 1024");
    if (e->md_block != threads) {
        e->md_block = threads;
        e->tstep = -1.0;  // rebuilt at the next step
    }
    return MPCD_OK;
}

int mpcd_ensemble_autotune(mpcd_ensemble *e, double tstep, int steps, int *md_wide, int *md_block, double *ms,
                           double *ms_each)
{
    if (!e || !md_wide || !md_block || !ms || steps < 1 || !(tstep > 0))
        return fail(MPCD_ERR_ARG, "bad ensemble autotune arguments");
    // the candidates: the wide MD path at each workgroup size, then a
    // workgroup per swimmer; each is one warm-up step (tables and graph
    // built) and `steps` timed steps of the run
    const int nc = 4;
    const int wide[nc] = {1, 1, 1, 0}, block[nc] = {kMdWideBlocks[0], kMdWideBlocks[1], kMdWideBlocks[2], kMdWideBlock};
    hipEvent_t t0 = nullptr, t1 = nullptr;
    int best = -1, rc = MPCD_OK;
    if (hipEventCreate(&t0) != hipSuccess || hipEventCreate(&t1) != hipSuccess)
        rc = fail(MPCD_ERR_HIP, "hipEventCreate failed");
    double best_ms = 0.0;
    for (int i = 0; i < nc && !rc; ++i) {
        if ((rc = mpcd_ensemble_set_md_wide(e, wide[i])) || (rc = mpcd_ensemble_set_md_block(e, block[i])) ||
            (rc = mpcd_ensemble_step(e, tstep, 1)))
            break;
        if (hipEventRecord(t0, e->stream) != hipSuccess) {
            rc = fail(MPCD_ERR_HIP, "hipEventRecord failed");
            break;
        }
        if ((rc = mpcd_ensemble_step(e, tstep, steps)))
            break;
        float t = 0.0f;
        if (hipEventRecord(t1, e->stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess ||
            hipEventElapsedTime(&t, t0, t1) != hipSuccess) {
            rc = fail(MPCD_ERR_HIP, "ensemble timing failed");
            break;
        }
        const double each = (double)t / steps;
        if (ms_each)
            ms_each[i] = each;
        if (best < 0 || each < best_ms) {
            best = i;
            best_ms = each;
        }
    }
    if (t0)
        (void)hipEventDestroy(t0);
    if (t1)
        (void)hipEventDestroy(t1);
    if (rc)
        return rc;
    if ((rc = mpcd_ensemble_set_md_wide(e, wide[best])) || (rc = mpcd_ensemble_set_md_block(e, block[best])))
        return rc;
    *md_wide = wide[best];
    *md_block = block[best];
    *ms = best_ms;
    return MPCD_OK;
}

int mpcd_ensemble_set_graph(mpcd_ensemble *e, int on)
{
    if (!e)
        return fail(MPCD_ERR_ARG, "null ensemble");
    if (e->use_graph != (on != 0)) {
        e->use_graph = on != 0;
        e->tstep = -1.0;  // rebuilt at the next step
    }
    return MPCD_OK;
}

int mpcd_ensemble_step(mpcd_ensemble *e, double tstep, int nsteps)
{
    if (!e)
        return fail(MPCD_ERR_ARG, "null ensemble");
    if (!(tstep > 0) || nsteps < 0)
        return fail(MPCD_ERR_ARG, "tstep must be positive and nsteps non-negative");
    for (int i = 0; i < nsteps; ++i) {
        // tables for this step length; the graph of the refreshes this step
        // makes (each union of the swimmers' schedules has a graph of its
        // own; the refresh kernel still checks each swimmer's own schedule;
        // a workgroup per swimmer takes its schedule on the device)
        if (tstep != e->tstep) {
            HIPCHK(hipStreamSynchronize(e->stream));
            if (int rc = ens_build(e, tstep)) {
                e->tstep = -1.0;
                return rc;
            }
        }
        const MdSchedule u = e->md_wide ? ens_schedule(e, tstep) : MdSchedule{};
        if (e->use_graph) {
            hipGraphExec_t x = nullptr;
            HIPCHK(e->graphs.get(u, e->wide.nmax,
                                 [&](const MdSchedule &use, hipGraph_t &g, hipGraphExec_t &ex) {
                                     return ens_capture(e, use, g, ex);
                                 },
                                 x));
            HIPCHK(hipGraphLaunch(x, e->stream));
        } else if (int rc = ens_enqueue(e, u)) {
            return rc;
        }
        for (mpcd_ctx *c : e->box)
            ens_advance_host(c, tstep);
    }
    return MPCD_OK;
}

int mpcd_ensemble_sync(mpcd_ensemble *e)
{
    if (!e)
        return fail(MPCD_ERR_ARG, "null ensemble");
    HIPCHK(hipStreamSynchronize(e->stream));
    return MPCD_OK;
}

int mpcd_ensemble_destroy(mpcd_ensemble *e)
{
    if (!e)
        return MPCD_OK;
    (void)hipStreamSynchronize(e->stream);
    for (size_t b = 0; b < e->box.size(); ++b) {
        e->box[b]->stream = e->own_stream[b];
        e->box[b]->ens = nullptr;
    }
    e->graphs.clear();
    (void)hipFree(e->d_boxes);
    (void)hipFree(e->d_args);
    (void)hipFree(e->d_tile_start);
    (void)hipFree(e->d_md);
    (void)hipFree(e->d_mdblk);
    (void)hipStreamDestroy(e->stream);
    delete e;
    return MPCD_OK;
}

}  // extern "C"
