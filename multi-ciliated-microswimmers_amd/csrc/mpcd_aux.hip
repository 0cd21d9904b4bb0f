// Kernels around the step: upload/binning of a host state, device-side
// fluid creation, overflow fix-up, download, bead binning, fills.
#include "mpcd_aux.hpp"

#include <cstring>
#include <rocprim/rocprim.hpp>

namespace mcm {

__global__ void k_fill_i32(int *p, int v, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, long long n)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        p[i] = v;
}

static inline int grid_for(long long n, int block = 256, int cap = 65536)
{
    long long g = (n + block - 1) / block;
    if (g < 1)
        g = 1;
    return (int)(g > cap ? cap : g);
}

void fill_i32(int *p, int v, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(256), 0, s, p, v, n);
}

void fill_u64(void *p, unsigned long long v, long long n, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill_u64, dim3(grid_for(n)), dim3(256), 0, s, (unsigned long long *)p, v, n);
}

// ------------------------------------------------------------------ upload
// Fluid::create/load + add_impuls of every particle (fluid.cpp:95-97,
// 147-149): each particle goes to the home segment of the tile that holds
// its cell under the shift a.
// Particle i (position x, velocity v, id i) as a record of its cell under
// the shift a, appended to the home segment of its tile.
template <bool F32, bool TRACK>
__device__ __forceinline__ void home_record(const Geom &g, const double x[3], const double v[3], const double a[3],
                                            long long i, uint32_t *home, int *home_cnt, int *flags)
{
    constexpr int W = F32 ? (TRACK ? 8 : 7) : (TRACK ? 14 : 12);
    uint32_t rec[W];
    int c[3];
    for (int k = 0; k < 3; ++k) {
        const double s = x[k] + a[k];
        c[k] = box_axis(s, g.L[k], g.invL[k]);
        if constexpr (F32) {
            float o = (float)(s - floor(s));
            if (o >= 1.0f) {
                o = 0.0f;
                c[k] = c[k] + 1 == g.L[k] ? 0 : c[k] + 1;
            }
            rec[k] = fbits(o);
            rec[3 + k] = fbits(quant_v(v[k]));
        } else {
            dsplit(x[k], rec[2 * k], rec[2 * k + 1]);
            dsplit(v[k], rec[6 + 2 * k], rec[7 + 2 * k]);
        }
    }
    if constexpr (F32) {
        rec[6] = pack_cell(c[0], c[1], c[2]);
        if constexpr (TRACK)
            rec[7] = (uint32_t)i;
    } else if constexpr (TRACK) {
        rec[12] = (uint32_t)i;
        rec[13] = 0;
    }
    const int t = c[0] / g.T[0] + g.nt[0] * (c[1] / g.T[1] + g.nt[1] * (c[2] / g.T[2]));
    const int slot = atomicAdd(&home_cnt[t], 1);
    if (slot < g.cap_home) {
        uint32_t *dst = home + ((long long)t * g.cap_home + slot) * W;
        for (int w = 0; w < W; ++w)
            dst[w] = rec[w];
    } else {
        atomicOr(flags, kFlagHomeCap);
    }
}

template <bool F32, bool TRACK>
__global__ void k_upload(Geom g, const double *__restrict__ pos3, const double *__restrict__ vel4, long long n,
                         double a0, double a1, double a2, uint32_t *home, int *home_cnt, int *flags)
{
    const double a[3] = {a0, a1, a2};
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const double x[3] = {pos3[3 * i], pos3[3 * i + 1], pos3[3 * i + 2]};
        const double v[3] = {vel4[4 * i], vel4[4 * i + 1], vel4[4 * i + 2]};
        home_record<F32, TRACK>(g, x, v, a, i, home, home_cnt, flags);
    }
}

void launch_upload(const Geom &g, bool f32, bool track, const double *pos3, const double *vel4, long long n,
                   const double a[3], uint32_t *home, int *home_cnt, int *flags, hipStream_t s)
{
    const dim3 grid(grid_for(n)), block(256);
    if (f32 && track)
        hipLaunchKernelGGL((k_upload<true, true>), grid, block, 0, s, g, pos3, vel4, n, a[0], a[1], a[2], home, home_cnt, flags);
    else if (f32)
        hipLaunchKernelGGL((k_upload<true, false>), grid, block, 0, s, g, pos3, vel4, n, a[0], a[1], a[2], home, home_cnt, flags);
    else if (track)
        hipLaunchKernelGGL((k_upload<false, true>), grid, block, 0, s, g, pos3, vel4, n, a[0], a[1], a[2], home, home_cnt, flags);
    else
        hipLaunchKernelGGL((k_upload<false, false>), grid, block, 0, s, g, pos3, vel4, n, a[0], a[1], a[2], home, home_cnt, flags);
}

// ------------------------------------------------------------ fluid create
// Device form of Fluid::create (fluid.cpp:55-98): uniform positions in the
// box [0, ext) per axis, Maxwell-Boltzmann velocities at kT, mass 1, drawn from the Philox
// stream (counter = particle index); the mean velocity is removed after.
// The mean is summed in a fixed order (per thread, per wave, per block,
// then the blocks' partial sums), so a seed gives the same bits every run.
constexpr uint64_t kTagFluid = 0x464c5544ull;

// A 256-thread block's sum of the per-thread (sx, sy, sz), in a fixed order;
// thread 0 gets it.
__device__ __forceinline__ void block_sum3_fixed(double &sx, double &sy, double &sz)
{
    __shared__ double s_w[4][3];
    for (int off = 32; off > 0; off >>= 1) {
        sx += __shfl_down(sx, off);
        sy += __shfl_down(sy, off);
        sz += __shfl_down(sz, off);
    }
    const int wave = threadIdx.x >> 6;
    if (lane_id() == 0) {
        s_w[wave][0] = sx;
        s_w[wave][1] = sy;
        s_w[wave][2] = sz;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        sx = ((s_w[0][0] + s_w[1][0]) + s_w[2][0]) + s_w[3][0];
        sy = ((s_w[0][1] + s_w[1][1]) + s_w[2][1]) + s_w[3][1];
        sz = ((s_w[0][2] + s_w[1][2]) + s_w[2][2]) + s_w[3][2];
    }
}

// Particle i of the created fluid, before the mean velocity is removed:
// Philox blocks {i, 0, 0, 'FLUD'} (position) and {i, 1, 0, 'FLUD'}
// (Box-Muller velocity, sk = sqrt(kT)).
__device__ __forceinline__ void create_draw(long long i, double L0, double L1, double L2, double sk, uint64_t seed,
                                            double x[3], double v[3])
{
    const U64x4 p = philox4x64((uint64_t)i, 0, 0, kTagFluid, seed, kKeyHi);
    const U64x4 q = philox4x64((uint64_t)i, 1, 0, kTagFluid, seed, kKeyHi);
    x[0] = u64_open01(p.v[0]) * L0;
    x[1] = u64_open01(p.v[1]) * L1;
    x[2] = u64_open01(p.v[2]) * L2;
    const double r0 = sqrt(-2.0 * log(u64_open01(q.v[0])));
    const double r1 = sqrt(-2.0 * log(u64_open01(q.v[2])));
    const double t0 = 2.0 * kPi * u64_open01(q.v[1]);
    const double t1 = 2.0 * kPi * u64_open01(q.v[3]);
    v[0] = r0 * cos(t0) * sk;
    v[1] = r0 * sin(t0) * sk;
    v[2] = r1 * cos(t1) * sk;
}

// pos3 / vel4 null: only the blocks' partial sums of the velocities (the
// no-allocation creation draws every particle again when it writes it).
__global__ __launch_bounds__(256) void k_create_fluid(double L0, double L1, double L2, double kT, uint64_t seed,
                                                      long long n, double *pos3, double *vel4, double *part)
{
    double sx = 0.0, sy = 0.0, sz = 0.0;
    const double sk = sqrt(kT);
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        double x[3], v[3];
        create_draw(i, L0, L1, L2, sk, seed, x, v);
        if (pos3) {
            pos3[3 * i + 0] = x[0];
            pos3[3 * i + 1] = x[1];
            pos3[3 * i + 2] = x[2];
            vel4[4 * i + 0] = v[0];
            vel4[4 * i + 1] = v[1];
            vel4[4 * i + 2] = v[2];
            vel4[4 * i + 3] = 1.0;
        }
        sx += v[0];
        sy += v[1];
        sz += v[2];
    }
    block_sum3_fixed(sx, sy, sz);
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x + 0] = sx;
        part[3 * blockIdx.x + 1] = sy;
        part[3 * blockIdx.x + 2] = sz;
    }
}

// Particle i drawn again, its velocity less the mean (the arithmetic of
// k_remove_mean), written as a record into its tile's home segment (the
// arithmetic of k_upload): the same records as create + remove + upload.
template <bool F32, bool TRACK>
__global__ __launch_bounds__(256) void k_create_records(Geom g, double L0, double L1, double L2, double kT,
                                                        uint64_t seed, long long n, const double *sum3, double a0,
                                                        double a1, double a2, uint32_t *home, int *home_cnt,
                                                        int *flags)
{
    const double sk = sqrt(kT);
    const double a[3] = {a0, a1, a2};
    const double mx = sum3[0] / (double)n, my = sum3[1] / (double)n, mz = sum3[2] / (double)n;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        double x[3], v[3];
        create_draw(i, L0, L1, L2, sk, seed, x, v);
        v[0] -= mx;
        v[1] -= my;
        v[2] -= mz;
        home_record<F32, TRACK>(g, x, v, a, i, home, home_cnt, flags);
    }
}

// The blocks' partial sums into sum3 (one block; fixed order).
__global__ __launch_bounds__(256) void k_sum_partials(const double *part, int nblk, double *sum3)
{
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int j = threadIdx.x; j < nblk; j += 256) {
        sx += part[3 * j + 0];
        sy += part[3 * j + 1];
        sz += part[3 * j + 2];
    }
    block_sum3_fixed(sx, sy, sz);
    if (threadIdx.x == 0) {
        sum3[0] = sx;
        sum3[1] = sy;
        sum3[2] = sz;
    }
}

__global__ void k_remove_mean(long long n, double *vel4, const double *sum3)
{
    const double mx = sum3[0] / (double)n, my = sum3[1] / (double)n, mz = sum3[2] / (double)n;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        vel4[4 * i + 0] -= mx;
        vel4[4 * i + 1] -= my;
        vel4[4 * i + 2] -= mz;
    }
}

void launch_create_records(const Geom &g, bool f32, bool track, const double ext[3], double kT, uint64_t seed,
                           long long n, double *sum3, const double a[3], uint32_t *home, int *home_cnt, int *flags,
                           hipStream_t s)
{
    const int nblk = grid_for(n, 256, kCreateBlocks);
    double *part = sum3 + 4;
    hipLaunchKernelGGL(k_create_fluid, dim3(nblk), dim3(256), 0, s, ext[0], ext[1], ext[2], kT, seed, n,
                       (double *)nullptr, (double *)nullptr, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, (const double *)part, nblk, sum3);
    const dim3 grid(grid_for(n)), block(256);
    if (f32 && track)
        hipLaunchKernelGGL((k_create_records<true, true>), grid, block, 0, s, g, ext[0], ext[1], ext[2], kT, seed, n,
                           (const double *)sum3, a[0], a[1], a[2], home, home_cnt, flags);
    else if (f32)
        hipLaunchKernelGGL((k_create_records<true, false>), grid, block, 0, s, g, ext[0], ext[1], ext[2], kT, seed, n,
                           (const double *)sum3, a[0], a[1], a[2], home, home_cnt, flags);
    else if (track)
        hipLaunchKernelGGL((k_create_records<false, true>), grid, block, 0, s, g, ext[0], ext[1], ext[2], kT, seed, n,
                           (const double *)sum3, a[0], a[1], a[2], home, home_cnt, flags);
    else
        hipLaunchKernelGGL((k_create_records<false, false>), grid, block, 0, s, g, ext[0], ext[1], ext[2], kT, seed,
                           n, (const double *)sum3, a[0], a[1], a[2], home, home_cnt, flags);
}

void launch_create_fluid(const double ext[3], double kT, uint64_t seed, long long n, double *pos3, double *vel4,
                         double *sum3, hipStream_t s)
{
    const int nblk = grid_for(n, 256, kCreateBlocks);
    double *part = sum3 + 4;
    hipLaunchKernelGGL(k_create_fluid, dim3(nblk), dim3(256), 0, s, ext[0], ext[1], ext[2], kT, seed, n, pos3, vel4,
                       part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, (const double *)part, nblk, sum3);
    hipLaunchKernelGGL(k_remove_mean, dim3(grid_for(n)), dim3(256), 0, s, n, vel4, (const double *)sum3);
}

// ------------------------------------------------------------ rigid sphere
// After the step: the sphere moved with its start-of-step velocity during
// the stream, x += v h (wrapped into the box); then v += dP / M,
// w += dL / I; the impulse is kept as the last step's and cleared.
__device__ void sphere_finish(SphereDev *sp, double h, const int (&L)[3])
{
    for (int k = 0; k < 3; ++k) {
        sp->X[k] += sp->V[k] * h;
        sp->X[k] -= L[k] * floor(sp->X[k] / L[k]);
        sp->V[k] += sp->dP[k] / sp->Ms;
        sp->W[k] += sp->dL[k] / sp->Is;
        sp->lastP[k] = sp->dP[k];
        sp->lastL[k] = sp->dL[k];
        sp->dP[k] = 0.0;
        sp->dL[k] = 0.0;
    }
    sp->h_last = h;
}

__global__ void k_sphere_finish(SphereDev *sp, double h, int L0, int L1, int L2)
{
    if (threadIdx.x != 0)
        return;
    const int L[3] = {L0, L1, L2};
    sphere_finish(sp, h, L);
}

void launch_sphere_finish(SphereDev *sp, double h, const int L[3], hipStream_t s)
{
    hipLaunchKernelGGL(k_sphere_finish, dim3(1), dim3(64), 0, s, sp, h, L[0], L[1], L[2]);
}

// --------------------------------------------------------- overflow fix-up
// Particles the step kernel could not place (full outbox or a move of more
// than one tile) go to the home segment of their tile here.
template <bool F32, bool TRACK>
__device__ __forceinline__ void fix_overflow_grid(const Geom &g, const uint32_t *ovf, const int *ovf_cnt,
                                                  int *ovf_cnt_next, double an0, double an1, double an2,
                                                  uint32_t *home, int *home_cnt, int *flags)
{
    constexpr int W = F32 ? (TRACK ? 8 : 7) : (TRACK ? 14 : 12);
    if (blockIdx.x == 0 && threadIdx.x == 0)
        *ovf_cnt_next = 0;
    const int n = min(*ovf_cnt, g.cap_ovf);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t *r = ovf + (long long)i * W;
        int c[3];
        if constexpr (F32) {
            c[0] = (int)(r[6] & 1023u);
            c[1] = (int)((r[6] >> 10) & 1023u);
            c[2] = (int)(r[6] >> 20);
        } else {
            const double an[3] = {an0, an1, an2};
            for (int k = 0; k < 3; ++k)
                c[k] = box_axis(djoin(r[2 * k], r[2 * k + 1]) + an[k], g.L[k], g.invL[k]);
        }
        const int t = c[0] / g.T[0] + g.nt[0] * (c[1] / g.T[1] + g.nt[1] * (c[2] / g.T[2]));
        const int slot = atomicAdd(&home_cnt[t], 1);
        if (slot < g.cap_home) {
            uint32_t *dst = home + ((long long)t * g.cap_home + slot) * W;
            for (int w = 0; w < W; ++w)
                dst[w] = r[w];
        } else {
            atomicOr(flags, kFlagHomeCap);
        }
    }
}

template <bool F32, bool TRACK>
__global__ void k_fix_overflow(Geom g, const uint32_t *ovf, const int *ovf_cnt, int *ovf_cnt_next, double an0,
                               double an1, double an2, uint32_t *home, int *home_cnt, int *flags)
{
    fix_overflow_grid<F32, TRACK>(g, ovf, ovf_cnt, ovf_cnt_next, an0, an1, an2, home, home_cnt, flags);
}

void launch_fix_overflow(const Geom &g, bool f32, bool track, const uint32_t *ovf, const int *ovf_cnt,
                         int *ovf_cnt_next, const double an[3], uint32_t *home, int *home_cnt, int *flags,
                         hipStream_t s)
{
    const dim3 grid(64), block(256);
    if (f32 && track)
        hipLaunchKernelGGL((k_fix_overflow<true, true>), grid, block, 0, s, g, ovf, ovf_cnt, ovf_cnt_next, an[0], an[1], an[2], home, home_cnt, flags);
    else if (f32)
        hipLaunchKernelGGL((k_fix_overflow<true, false>), grid, block, 0, s, g, ovf, ovf_cnt, ovf_cnt_next, an[0], an[1], an[2], home, home_cnt, flags);
    else if (track)
        hipLaunchKernelGGL((k_fix_overflow<false, true>), grid, block, 0, s, g, ovf, ovf_cnt, ovf_cnt_next, an[0], an[1], an[2], home, home_cnt, flags);
    else
        hipLaunchKernelGGL((k_fix_overflow<false, false>), grid, block, 0, s, g, ovf, ovf_cnt, ovf_cnt_next, an[0], an[1], an[2], home, home_cnt, flags);
}

// ---------------------------------------------------------------- download
__global__ void k_tile_totals(Geom g, const int *home_cnt, const int *ob_cnt, int *tot)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.ntiles)
        return;
    const TileCoord tc = tile_coord(g, t);
    int sum = 0;
    for (int s = 0; s < 27; ++s) {
        int cnt;
        long long base;
        segment_info(g, tc, s, home_cnt, ob_cnt, cnt, base);
        sum += cnt;
    }
    tot[t] = sum;
}

void launch_tile_totals(const Geom &g, const int *home_cnt, const int *ob_cnt, int *tot, hipStream_t s)
{
    hipLaunchKernelGGL(k_tile_totals, dim3((g.ntiles + 255) / 256), dim3(256), 0, s, g, home_cnt, ob_cnt, tot);
}

// One block per tile; particle j of the tile goes to out index id (with ids)
// or prefix[t] + j.  FP32-bulk positions decode as cell - a + offset.
template <bool F32, bool TRACK>
__global__ void k_download(Geom g, const uint32_t *home, const uint32_t *ob, const int *home_cnt, const int *ob_cnt,
                           const long long *prefix, double a0, double a1, double a2, double *pos3, double *vel4)
{
    constexpr int W = F32 ? (TRACK ? 8 : 7) : (TRACK ? 14 : 12);
    __shared__ int s_seg[28];
    __shared__ long long s_base[27];
    const TileCoord tc = tile_coord(g, blockIdx.x);
    if (threadIdx.x < 27) {
        int cnt;
        long long base;
        segment_info(g, tc, threadIdx.x, home_cnt, ob_cnt, cnt, base);
        s_seg[threadIdx.x + 1] = cnt;
        s_base[threadIdx.x] = base;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        s_seg[0] = 0;
        for (int s = 1; s <= 27; ++s)
            s_seg[s] += s_seg[s - 1];
    }
    __syncthreads();
    const int total = s_seg[27];
    for (int j = threadIdx.x; j < total; j += blockDim.x) {
        int s = 0;
        while (j >= s_seg[s + 1])
            ++s;
        const uint32_t *r = (s == 0 ? home : ob) + (s_base[s] + (j - s_seg[s])) * W;
        long long o;
        if constexpr (TRACK)
            o = (long long)r[F32 ? 7 : 12];
        else
            o = prefix[blockIdx.x] + j;
        if constexpr (F32) {
            const int c[3] = {(int)(r[6] & 1023u), (int)((r[6] >> 10) & 1023u), (int)(r[6] >> 20)};
            const double a[3] = {a0, a1, a2};
            for (int k = 0; k < 3; ++k) {
                pos3[3 * o + k] = (double)c[k] - a[k] + (double)bitsf(r[k]);
                vel4[4 * o + k] = (double)bitsf(r[3 + k]);
            }
        } else {
            for (int k = 0; k < 3; ++k) {
                pos3[3 * o + k] = djoin(r[2 * k], r[2 * k + 1]);
                vel4[4 * o + k] = djoin(r[6 + 2 * k], r[7 + 2 * k]);
            }
        }
        vel4[4 * o + 3] = 1.0;
    }
}

// Raw records (checkpoints): record j of tile t goes to words[(prefix[t] + j) W].
template <int W>
__global__ void k_download_raw(Geom g, const uint32_t *home, const uint32_t *ob, const int *home_cnt,
                               const int *ob_cnt, const long long *prefix, uint32_t *words)
{
    __shared__ int s_seg[28];
    __shared__ long long s_base[27];
    const TileCoord tc = tile_coord(g, blockIdx.x);
    if (threadIdx.x < 27) {
        int cnt;
        long long base;
        segment_info(g, tc, threadIdx.x, home_cnt, ob_cnt, cnt, base);
        s_seg[threadIdx.x + 1] = cnt;
        s_base[threadIdx.x] = base;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        s_seg[0] = 0;
        for (int s = 1; s <= 27; ++s)
            s_seg[s] += s_seg[s - 1];
    }
    __syncthreads();
    const int total = s_seg[27];
    for (int j = threadIdx.x; j < total; j += blockDim.x) {
        int s = 0;
        while (j >= s_seg[s + 1])
            ++s;
        const uint32_t *r = (s == 0 ? home : ob) + (s_base[s] + (j - s_seg[s])) * W;
        uint32_t *o = words + (prefix[blockIdx.x] + j) * W;
        for (int w = 0; w < W; ++w)
            o[w] = r[w];
    }
}

// Raw records back into the home segments of the tiles that hold their
// cells under the shift a (FP32-bulk: the cell key; FP64: the position).
template <bool F32, int W>
__global__ void k_upload_raw(Geom g, const uint32_t *words, long long n, double a0, double a1, double a2,
                             uint32_t *home, int *home_cnt, int *flags)
{
    const double a[3] = {a0, a1, a2};
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
        const uint32_t *r = words + i * W;
        int c[3];
        if constexpr (F32) {
            c[0] = (int)(r[6] & 1023u);
            c[1] = (int)((r[6] >> 10) & 1023u);
            c[2] = (int)(r[6] >> 20);
        } else {
            for (int k = 0; k < 3; ++k)
                c[k] = box_axis(djoin(r[2 * k], r[2 * k + 1]) + a[k], g.L[k], g.invL[k]);
        }
        const int t = c[0] / g.T[0] + g.nt[0] * (c[1] / g.T[1] + g.nt[1] * (c[2] / g.T[2]));
        const int slot = atomicAdd(&home_cnt[t], 1);
        if (slot < g.cap_home) {
            uint32_t *dst = home + ((long long)t * g.cap_home + slot) * W;
            for (int w = 0; w < W; ++w)
                dst[w] = r[w];
        } else {
            atomicOr(flags, kFlagHomeCap);
        }
    }
}

void launch_download_raw(const Geom &g, bool f32, bool track, const uint32_t *home, const uint32_t *ob,
                         const int *home_cnt, const int *ob_cnt, const long long *prefix, uint32_t *words,
                         hipStream_t s)
{
    const dim3 grid(g.ntiles), block(256);
    const int W = f32 ? (track ? 8 : 7) : (track ? 14 : 12);
    if (W == 7)
        hipLaunchKernelGGL(k_download_raw<7>, grid, block, 0, s, g, home, ob, home_cnt, ob_cnt, prefix, words);
    else if (W == 8)
        hipLaunchKernelGGL(k_download_raw<8>, grid, block, 0, s, g, home, ob, home_cnt, ob_cnt, prefix, words);
    else if (W == 12)
        hipLaunchKernelGGL(k_download_raw<12>, grid, block, 0, s, g, home, ob, home_cnt, ob_cnt, prefix, words);
    else
        hipLaunchKernelGGL(k_download_raw<14>, grid, block, 0, s, g, home, ob, home_cnt, ob_cnt, prefix, words);
}

void launch_upload_raw(const Geom &g, bool f32, bool track, const uint32_t *words, long long n, const double a[3],
                       uint32_t *home, int *home_cnt, int *flags, hipStream_t s)
{
    const dim3 grid(grid_for(n)), block(256);
    if (f32 && track)
        hipLaunchKernelGGL((k_upload_raw<true, 8>), grid, block, 0, s, g, words, n, a[0], a[1], a[2], home, home_cnt, flags);
    else if (f32)
        hipLaunchKernelGGL((k_upload_raw<true, 7>), grid, block, 0, s, g, words, n, a[0], a[1], a[2], home, home_cnt, flags);
    else if (track)
        hipLaunchKernelGGL((k_upload_raw<false, 14>), grid, block, 0, s, g, words, n, a[0], a[1], a[2], home, home_cnt, flags);
    else
        hipLaunchKernelGGL((k_upload_raw<false, 12>), grid, block, 0, s, g, words, n, a[0], a[1], a[2], home, home_cnt, flags);
}

void launch_download(const Geom &g, bool f32, bool track, const uint32_t *home, const uint32_t *ob, const int *home_cnt,
                     const int *ob_cnt, const long long *prefix, const double a[3], double *pos3, double *vel4,
                     hipStream_t s)
{
    const dim3 grid(g.ntiles), block(256);
    if (f32 && track)
        hipLaunchKernelGGL((k_download<true, true>), grid, block, 0, s, g, home, ob, home_cnt, ob_cnt, prefix, a[0], a[1], a[2], pos3, vel4);
    else if (f32)
        hipLaunchKernelGGL((k_download<true, false>), grid, block, 0, s, g, home, ob, home_cnt, ob_cnt, prefix, a[0], a[1], a[2], pos3, vel4);
    else if (track)
        hipLaunchKernelGGL((k_download<false, true>), grid, block, 0, s, g, home, ob, home_cnt, ob_cnt, prefix, a[0], a[1], a[2], pos3, vel4);
    else
        hipLaunchKernelGGL((k_download<false, false>), grid, block, 0, s, g, home, ob, home_cnt, ob_cnt, prefix, a[0], a[1], a[2], pos3, vel4);
}

// ------------------------------------------------------------------- beads
// Bead binning: key = tile * NC + local cell under the shift a; a stable
// radix sort keeps bead index order inside each cell (the order beads join
// a cell in the reference: md_objects order, then rod, then bead).
// Sort keys of the beads: tile * nc + local cell; a fixed bead (mobile[i]
// == 0) takes kBeadFixedKey, sorts after every tile and joins no collision
// (the reference adds only beads j >= startidx, mdobject.cpp:62-72,153-160).
constexpr unsigned kBeadFixedKey = 0xffffffffu;

__global__ void k_bead_keys(Geom g, int nc, const double *bpos3, const unsigned char *mobile, int nb, double a0,
                            double a1, double a2, unsigned int *keys, unsigned int *vals)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb)
        return;
    if (mobile && !mobile[i]) {
        keys[i] = kBeadFixedKey;
        vals[i] = (unsigned)i;
        return;
    }
    const double a[3] = {a0, a1, a2};
    int c[3], l[3], t[3];
    for (int k = 0; k < 3; ++k) {
        c[k] = box_axis(bpos3[3 * i + k] + a[k], g.L[k], g.invL[k]);
        t[k] = c[k] / g.T[k];
        l[k] = c[k] - t[k] * g.T[k];
    }
    const int tile = t[0] + g.nt[0] * (t[1] + g.nt[1] * t[2]);
    const int lc = l[0] + g.T[0] * (l[1] + g.T[1] * l[2]);
    keys[i] = (unsigned)tile * (unsigned)nc + (unsigned)lc;
    vals[i] = (unsigned)i;
}

__global__ void k_bead_gather(int nc, int ntiles, int nb, const unsigned *skeys, const unsigned *svals,
                              const double *bvel4, int *b_lc, int *b_idx, double *b_vel4, int *tile_start)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb)
        return;
    // fixed beads (sorted last) join no tile
    const unsigned key = skeys[i];
    if (key != kBeadFixedKey) {
        b_lc[i] = (int)(key % (unsigned)nc);
        const int src = (int)svals[i];
        b_idx[i] = src;
        for (int k = 0; k < 4; ++k)
            b_vel4[4 * (long long)i + k] = bvel4[4 * (long long)src + k];
    }
}

// First sorted bead of each tile t (and of t = ntiles: the count of beads
// that join a tile): the first key at or above t nc, by binary search.
__device__ __forceinline__ int first_key_at_least(const unsigned *skeys, int nb, unsigned key)
{
    int lo = 0, hi = nb;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (skeys[mid] < key)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

__global__ void k_bead_tile_start(int nc, int ntiles, int nb, const unsigned *skeys, int *tile_start)
{
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t <= ntiles; t += gridDim.x * blockDim.x)
        tile_start[t] = first_key_at_least(skeys, nb, (unsigned)t * (unsigned)nc);
}

size_t bead_sort_temp_bytes(int nb_max)
{
    size_t bytes = 0;
    rocprim::radix_sort_pairs(nullptr, bytes, (unsigned *)nullptr, (unsigned *)nullptr,
                                                  (unsigned *)nullptr, (unsigned *)nullptr, (unsigned)nb_max);
    return bytes;
}

int launch_bead_bin(const Geom &g, int nc, const double *bpos3, const double *bvel4, const unsigned char *mobile,
                    int nb, const double a[3], unsigned *keys, unsigned *vals, unsigned *skeys, unsigned *svals,
                    void *temp, size_t temp_bytes, int *b_lc, int *b_idx, double *b_vel4, int *tile_start,
                    hipStream_t s)
{
    if (nb <= 0) {
        fill_i32(tile_start, 0, g.ntiles + 1, s);
        return 0;
    }
    hipLaunchKernelGGL(k_bead_keys, dim3((nb + 255) / 256), dim3(256), 0, s, g, nc, bpos3, mobile, nb, a[0], a[1],
                       a[2], keys, vals);
    size_t bytes = temp_bytes;
    if (rocprim::radix_sort_pairs(temp, bytes, keys, skeys, vals, svals, (unsigned)nb, 0, 32, s) != hipSuccess)
        return -1;
    hipLaunchKernelGGL(k_bead_gather, dim3((nb + 255) / 256), dim3(256), 0, s, nc, g.ntiles, nb, skeys, svals, bvel4,
                       b_lc, b_idx, b_vel4, tile_start);
    hipLaunchKernelGGL(k_bead_tile_start, dim3(grid_for((long long)g.ntiles + 1)), dim3(256), 0, s, nc, g.ntiles, nb,
                       (const unsigned *)skeys, tile_start);
    return 0;
}

// ------------------------------------------------------------- diagnostics
// Folds the per-step slots into the running totals (fixed summation order)
// and clears them; with reset_run the dissipation totals restart from zero.
// Called by every thread of a block of at least kDiagSlots threads.
__device__ __forceinline__ void diag_fold_block(Diag *d, int reset_run)
{
    __shared__ long long s_q[kDiagSlots];
    __shared__ long long s_rv[kDiagSlots];
    __shared__ double s_d[kDiagSlots];
    const int i = threadIdx.x;
    const bool slot_thread = i < kDiagSlots;
    if (slot_thread) {
        const DiagSlot &sl = d->slot[i];
        s_q[i] = sl.q_diss;
        s_rv[i] = sl.rel_viol;
        s_d[i] = sl.diss;
    }
    __syncthreads();
    for (int off = kDiagSlots / 2; off > 0; off >>= 1) {
        if (i < off) {
            s_q[i] += s_q[i + off];
            s_rv[i] += s_rv[i + off];
            s_d[i] += s_d[i + off];
        }
        __syncthreads();
    }
    if (!slot_thread)
        return;
    DiagSlot &sl = d->slot[i];
    if (i == 0) {
        if (reset_run) {
            d->run_q_diss = 0;
            d->run_diss = 0.0;
        } else {
            d->run_q_diss += s_q[0];
            d->run_diss += s_d[0];
            d->run_rel_viol += s_rv[0];
            d->defer_n = 0;
            d->bead_n = 0;
        }
    }
    sl.q_diss = 0;
    sl.diss = 0.0;
    if (!reset_run) {
        sl.rel_viol = 0;
        sl.q_mom[0] = sl.q_mom[1] = sl.q_mom[2] = 0;
        sl.mom[0] = sl.mom[1] = sl.mom[2] = 0.0;
        sl.e2 = 0.0;
        sl.n = 0;
        sl.max_load = 0;
    }
}

__global__ __launch_bounds__(kDiagSlots) void k_begin_step(Diag *d, int reset_run)
{
    diag_fold_block(d, reset_run);
}

// --------------------------------------------------------------- ensembles
// Per box, before its step: the diagnostics fold, this step's arguments
// (the template of the current buffer parity with the shift, the next
// shift drawn as the single-box host path draws it, the step counter and
// the overflow counter of the step's parity) and, for a box with a swimmer,
// the bead binning of k_bead_keys + radix sort + k_bead_gather, done here
// by one workgroup: (key, bead index) pairs sorted by a bitonic network in
// LDS, which orders them exactly as the stable sort of the keys does.
constexpr int kEnsBlock = 1024;
constexpr uint64_t kTagShift = 0x53484654ull;  // the host's shift_of_step stream

__device__ __forceinline__ double shift_draw(uint64_t seed, uint64_t step, int k)
{
    const U64x4 o = philox4x64(0, step, 0, kTagShift, seed, kKeyHi);
    return (double)(1 + (o.v[k] >> 40) % 16777215ull) * (1.0 / 16777216.0);
}

__global__ __launch_bounds__(kEnsBlock) void k_ens_begin(EnsBox *__restrict__ boxes, StepArgs *__restrict__ args)
{
    __shared__ unsigned long long s_key[kEnsMaxBeads];
    EnsBox &B = boxes[blockIdx.x];
    const StepArgs &T = B.tmpl[B.cur];
    diag_fold_block(T.diag, 0);
    if (threadIdx.x == 0) {
        StepArgs A = T;
        A.step = B.step;
        A.ovf_cnt = B.ovf_cnt2 + (int)(B.step & 1);
        // the swimmer's positions at the collision, by bead index (MPC-AT+a's
        // lever arms of the beads read them)
        A.b_pos3 = B.md ? B.md->x[B.md->cur] : nullptr;
        for (int k = 0; k < 3; ++k) {
            A.a[k] = B.a[k];
            A.a_next[k] = shift_draw(T.seed, B.step + 1, k);
            A.da[k] = (float)(A.a_next[k] - A.a[k]);
        }
        args[blockIdx.x] = A;
    }
    const int nb = T.nb;
    if (nb == 0)
        return;
    const Geom &g = T.g;
    const int nc = g.T[0] * g.T[1] * g.T[2];
    const double *bx = B.md->x[B.md->cur];
    int P = 64;
    while (P < nb)
        P <<= 1;
    for (int i = threadIdx.x; i < P; i += kEnsBlock) {
        unsigned long long key = ~0ull;
        if (i < nb && B.md->M.mobile && !B.md->M.mobile[i]) {  // fixed: after every tile, in index order
            key = ((unsigned long long)kBeadFixedKey << 32) | (unsigned)i;
        } else if (i < nb) {
            int c[3], l[3], t[3];
            for (int k = 0; k < 3; ++k) {
                c[k] = box_axis(bx[3 * i + k] + B.a[k], g.L[k], g.invL[k]);
                t[k] = c[k] / g.T[k];
                l[k] = c[k] - t[k] * g.T[k];
            }
            const int tile = t[0] + g.nt[0] * (t[1] + g.nt[1] * t[2]);
            const int lc = l[0] + g.T[0] * (l[1] + g.T[1] * l[2]);
            key = ((unsigned long long)((unsigned)tile * (unsigned)nc + (unsigned)lc) << 32) | (unsigned)i;
        }
        s_key[i] = key;
    }
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < P; i += kEnsBlock) {
                const int p = i ^ j;
                if (p > i) {
                    const unsigned long long x = s_key[i], y = s_key[p];
                    if ((x > y) == ((i & k) == 0)) {
                        s_key[i] = y;
                        s_key[p] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    const double *bv = B.md->M.v;
    int *b_lc = const_cast<int *>(T.b_lc), *b_idx = const_cast<int *>(T.b_idx);
    int *tile_start = const_cast<int *>(T.b_tile_start);
    double *b_vel4 = const_cast<double *>(T.b_vel4);
    for (int i = threadIdx.x; i < nb; i += kEnsBlock) {
        const unsigned key = (unsigned)(s_key[i] >> 32);
        if (key != kBeadFixedKey) {
            const int src = (int)(unsigned)s_key[i];
            b_lc[i] = (int)(key % (unsigned)nc);
            b_idx[i] = src;
            for (int k = 0; k < 4; ++k)
                b_vel4[4 * (long long)i + k] = bv[4 * (long long)src + k];
        }
    }
    // first sorted bead of each tile (the first key at or above t nc)
    for (int t = threadIdx.x; t <= g.ntiles; t += kEnsBlock) {
        const unsigned long long want = (unsigned long long)((unsigned)t * (unsigned)nc) << 32;
        int lo = 0, hi = nb;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (s_key[mid] < want)
                lo = mid + 1;
            else
                hi = mid;
        }
        tile_start[t] = lo;
    }
}

// Per box, after its step: the overflow fix-up (blockIdx.y = box), and the
// box's running state advanced: shift, step counter, buffer parity.
template <bool F32, bool TRACK>
__global__ void k_ens_fix(EnsBox *__restrict__ boxes, const StepArgs *__restrict__ args)
{
    EnsBox &B = boxes[blockIdx.y];
    const StepArgs &A = args[blockIdx.y];
    fix_overflow_grid<F32, TRACK>(A.g, A.ovf, A.ovf_cnt, B.ovf_cnt2 + (int)(1 - (A.step & 1)), A.a_next[0],
                                  A.a_next[1], A.a_next[2], A.out_home, A.out_home_cnt, A.flags);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        for (int k = 0; k < 3; ++k)
            B.a[k] = A.a_next[k];
        B.step = A.step + 1;
        B.cur ^= 1;
        if (A.sphere)  // the step's impulses move the box's rigid sphere
            sphere_finish(A.sphere, A.h, A.g.L);
    }
}

void launch_ens_begin(EnsBox *boxes, StepArgs *args, int nbox, hipStream_t s)
{
    hipLaunchKernelGGL(k_ens_begin, dim3(nbox), dim3(kEnsBlock), 0, s, boxes, args);
}

void launch_ens_fix(EnsBox *boxes, const StepArgs *args, int nbox, bool f32, bool track, hipStream_t s)
{
    const dim3 grid(64, nbox), block(256);
    if (f32 && track)
        hipLaunchKernelGGL((k_ens_fix<true, true>), grid, block, 0, s, boxes, args);
    else if (f32)
        hipLaunchKernelGGL((k_ens_fix<true, false>), grid, block, 0, s, boxes, args);
    else if (track)
        hipLaunchKernelGGL((k_ens_fix<false, true>), grid, block, 0, s, boxes, args);
    else
        hipLaunchKernelGGL((k_ens_fix<false, false>), grid, block, 0, s, boxes, args);
}

void launch_begin_step(Diag *d, bool reset_run, hipStream_t s)
{
    hipLaunchKernelGGL(k_begin_step, dim3(1), dim3(kDiagSlots), 0, s, d, reset_run ? 1 : 0);
}

}  // namespace mcm
