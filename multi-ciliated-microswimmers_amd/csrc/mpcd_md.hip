// Bead-spring MD sub-steps on gfx950 (see mpcd_md.hpp).
//
// All arithmetic is FP64 and, compiled with -ffp-contract=off, rounds each
// operation separately in the reference's order.  pow(x, 2.0) and
// pow(x, 2.5) of the beat rule are x*x and (x*x)*sqrt(x), what a
// -ffast-math build of the reference computes.
#include <cmath>

#include "mpcd_device.hpp"
#include "mpcd_md.hpp"

namespace mcm {

namespace {

constexpr int kMdBlock = 1024;
constexpr int kMdHeavy = 32;  // beads with more terms are summed by a whole wave (the sphere's centre)
constexpr int kMdWideTerms = 65536;  // above: one launch per phase over the whole GPU

__device__ __forceinline__ double md_sqdist(const double *a, const double *b)
{
    // Matrix::sqdistance (src/libcilia/Matrix.h:412-419)
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double d = a[i] - b[i];
        s += d * d;
    }
    return s;
}

// One force term: the contribution (x_self - x_partner) fac of
// MDObject::spring (src/libcilia/mdobject.cpp:8-17), or an attachment
// spring's (pos1 - pos0) fac of Clami::global_forces (src/clami.cpp:188-198),
// negated on its second bead (force1 -= F).  Adding the stored value to the
// bead's running sum rounds exactly as the reference's += / -= do.
__device__ __forceinline__ void md_term(const MdTerm &T, int self, const double *__restrict__ x,
                                        const MdCilium *__restrict__ cil, const double *__restrict__ newlen,
                                        double (&out)[3])
{
    const double *xs = x + 3 * (long long)self;
    const double *xp = x + 3 * (long long)T.partner;
    double rest = T.b;
    const int kind = T.flags & kTermKindMask;
    if (kind != kTermConst) {
        const MdCilium &C = cil[T.cil];
        const double nl = newlen[C.len_off + T.idx];
        rest = kind == kTermNewLen ? nl : sqrt(C.bondlength * C.bondlength + C.bondlength * nl);
    }
    if (T.flags & kTermAttach) {
        const bool second = (T.flags & kTermAttachB) != 0;
        const double *p1 = second ? xp : xs, *p0 = second ? xs : xp;
        const double d = sqrt(md_sqdist(p1, p0));
        double F[3] = {0.0, 0.0, 0.0};
        if (d != 0) {
            const double fac = -2.0 * T.k * (1.0 - rest / d);
#pragma unroll
            for (int i = 0; i < 3; ++i)
                F[i] = (p1[i] - p0[i]) * fac;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
            out[i] = second ? -F[i] : F[i];
    } else {
        const double fac = -2.0 * T.k * (1.0 - rest / sqrt(md_sqdist(xs, xp)));
#pragma unroll
        for (int i = 0; i < 3; ++i)
            out[i] = (xs[i] - xp[i]) * fac;
    }
}

// Cilia::length (src/cilia.cpp:80-121): the beat profile and the stall rule.
__device__ double cilium_length(const MdCilium &C, const double *x, const double *oldlen, int i)
{
    const double bl = C.bondlength;
    const double *x0 = x + 3 * (long long)C.first;  // rod 0
    double l;
    double tl = sqrt(md_sqdist(x0 + 3 * i, x0 + 3 * (i + 1)));
    if (i < C.midpnt - 2.0) {
        const double t = 1.0 - (i - C.num_anker - 1.0) / (C.rodlength - C.num_anker - 1.0);
        const double p25 = (t * t) * sqrt(t);
        l = bl + C.beatstrength * p25 * 0.5 * (1.0 - 1.0 / (C.midpnt - i - 1.0));
    } else {
        const double d = i - C.midpnt;
        l = bl - 2.0 * C.beatstrength / (d * d + 1.0);
    }
    if (fabs(l - tl) > C.bending) {
        tl = tl + C.bending - 2.0 * C.bending * (double)(l < tl);
        const double o = oldlen[C.len_off + i];
        if (((l - o) * (tl - o)) > 0)
            l = tl;
        else
            l = o;
    }
    return l;
}

// Cilia::localbend(0, rodlength) (src/cilia.cpp:68-78).
__device__ double cilium_bend(const MdCilium &C, const double *x)
{
    const int R = C.rodlength;
    const double *p = x + 3 * (long long)C.first;
    double l = 0;
    for (int i = 0; i < R - 1; i++) {
        l += 2.0 * sqrt(md_sqdist(p + 3 * i, p + 3 * (i + 1)));
        l -= sqrt(md_sqdist(p + 3 * (R + i), p + 3 * (R + i + 1)));
        l -= sqrt(md_sqdist(p + 3 * (2 * R + i), p + 3 * (2 * R + i + 1)));
    }
    return l;
}

// Kick (v += F c) and, for sub-steps, drift into the other buffer (with the
// walls' bounce-back); fixed beads stay put.  The bead's force is stored.
__device__ __forceinline__ void md_kick_drift(int b, const double f[3], double c, bool drift, double dt,
                                              const double *__restrict__ x, double *__restrict__ xn,
                                              double *__restrict__ v, double *__restrict__ fo,
                                              const unsigned char *__restrict__ mobile, int walls, double ztop,
                                              double vwx, double vwy)
{
    double *vb = v + 4 * (long long)b;
    const double *xs = x + 3 * (long long)b;
    double *xb = xn + 3 * (long long)b;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        fo[3 * (long long)b + i] = f[i];
    if (!mobile || mobile[b]) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
            vb[i] += f[i] * c;
        if (drift) {
#pragma unroll
            for (int i = 0; i < 3; ++i)
                xb[i] = xs[i] + vb[i] * dt;
            if (walls)
                wall_bounce(walls, ztop, vwx, vwy, xb[0], xb[1], xb[2], vb[0], vb[1], vb[2]);
        }
    } else if (drift) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
            xb[i] = xs[i];
    }
}

// ---- the phases of one force evaluation, shared by the one-workgroup
//      kernel and the wide (one launch per phase) path; every loop is a
//      stride loop over its items

// the beat's rest lengths (Cilia::recalcforce's refresh, cilia.cpp:206-212)
__device__ __forceinline__ void md_phase_refresh(int i0, int stride, int ncil, const MdCilium *__restrict__ cil,
                                                 const double *__restrict__ x, double *__restrict__ newlen,
                                                 double *__restrict__ oldlen)
{
    for (int c = i0; c < ncil; c += stride) {
        const MdCilium C = cil[c];
        for (int i = C.num_anker - 1; i < C.rodlength - 1; ++i) {
            const double l = cilium_length(C, x, oldlen, i);
            newlen[C.len_off + i] = l;
            oldlen[C.len_off + i] = l;
        }
    }
}

// every force term
__device__ __forceinline__ void md_phase_terms(int i0, int stride, int nterms, const MdTerm *__restrict__ terms,
                                               const int *__restrict__ tself, const double *__restrict__ x,
                                               const MdCilium *__restrict__ cil, const double *__restrict__ newlen,
                                               double *__restrict__ contrib)
{
    for (int t = i0; t < nterms; t += stride) {
        double cval[3];
        md_term(terms[t], tself[t], x, cil, newlen, cval);
        contrib[3 * (long long)t] = cval[0];
        contrib[3 * (long long)t + 1] = cval[1];
        contrib[3 * (long long)t + 2] = cval[2];
    }
}

struct MdKick {
    double c, dt, ztop, vwx, vwy;
    int drift, walls;
};

// each bead with a short term list: its ordered sum, kick, drift
__device__ __forceinline__ void md_phase_light(int i0, int stride, int nb, const int *__restrict__ tstart,
                                               const double *__restrict__ contrib, const double *__restrict__ x,
                                               double *__restrict__ xn, double *__restrict__ v,
                                               double *__restrict__ fo, const unsigned char *__restrict__ mobile,
                                               const MdKick &K)
{
    for (int b = i0; b < nb; b += stride) {
        double f[3] = {0.0, 0.0, 0.0};
        const int t1 = tstart[b + 1];
        int t = tstart[b];
        if (t1 - t > kMdHeavy)
            continue;  // summed by a wave
        // eight terms' loads in flight at a time, added in list order
        for (; t + 8 <= t1; t += 8) {
            double cv[24];
#pragma unroll
            for (int j = 0; j < 24; ++j)
                cv[j] = contrib[3 * (long long)t + j];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                f[0] += cv[3 * j];
                f[1] += cv[3 * j + 1];
                f[2] += cv[3 * j + 2];
            }
        }
        for (; t < t1; ++t) {
#pragma unroll
            for (int i = 0; i < 3; ++i)
                f[i] += contrib[3 * (long long)t + i];
        }
        md_kick_drift(b, f, K.c, K.drift, K.dt, x, xn, v, fo, mobile, K.walls, K.ztop, K.vwx, K.vwy);
    }
}

// a bead with a long term list, by one wave: 64 terms at a time staged in
// LDS (hv: two chunks, one coalesced load per lane, the next chunk loading
// while the current one is added), lane 0 adds them in list order
__device__ __forceinline__ void md_heavy_bead(int b, int lane, double (*hv)[64 * 3], const int *__restrict__ tstart,
                                              const double *__restrict__ contrib, const double *__restrict__ x,
                                              double *__restrict__ xn, double *__restrict__ v,
                                              double *__restrict__ fo, const unsigned char *__restrict__ mobile,
                                              const MdKick &K)
{
    const int t0 = tstart[b], t1 = tstart[b + 1];
    double f[3] = {0.0, 0.0, 0.0};
    auto stage = [&](int cb, int buf) {
        const int t = cb + lane;
        if (t < t1) {
#pragma unroll
            for (int i = 0; i < 3; ++i)
                hv[buf][3 * lane + i] = contrib[3 * (long long)t + i];
        }
    };
    auto wave_sync = []() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    stage(t0, 0);
    int buf = 0;
    for (int cb = t0; cb < t1; cb += 64, buf ^= 1) {
        wave_sync();
        if (cb + 64 < t1)
            stage(cb + 64, buf ^ 1);
        if (lane == 0) {
            const double *h = hv[buf];
            const int m = min(64, t1 - cb);
            int j = 0;
            for (; j + 8 <= m; j += 8) {
                double cv[24];
#pragma unroll
                for (int q = 0; q < 24; ++q)
                    cv[q] = h[3 * j + q];
#pragma unroll
                for (int q = 0; q < 8; ++q) {
                    f[0] += cv[3 * q];
                    f[1] += cv[3 * q + 1];
                    f[2] += cv[3 * q + 2];
                }
            }
            for (; j < m; ++j) {
                f[0] += h[3 * j];
                f[1] += h[3 * j + 1];
                f[2] += h[3 * j + 2];
            }
        }
    }
    if (lane == 0)
        md_kick_drift(b, f, K.c, K.drift, K.dt, x, xn, v, fo, mobile, K.walls, K.ztop, K.vwx, K.vwy);
    wave_sync();
}

// the beat state machine at the end of the step (ConstantBeatCilia.cpp:8-23,
// cilia.cpp:241-258)
__device__ __forceinline__ void md_phase_switch(int i0, int stride, int ncil, MdCilium *__restrict__ cil,
                                                const double *__restrict__ x, double g_end, double tstep)
{
    for (int c = i0; c < ncil; c += stride) {
        MdCilium C = cil[c];
        const double start = C.num_anker;
        if (C.constant_beat) {
            if (g_end - C.Tstroke >= C.power_stroke_duration && C.midpnt == start)
                C.midpnt += 1;
            if (g_end - C.Tstroke - C.power_stroke_duration >= C.recovery_stroke_duration && C.midpnt > start) {
                C.Tstroke = g_end;
                C.midpnt = start;
            }
        } else {
            if ((C.midpnt == start) && g_end - C.Tstroke > 100.0)
                C.midpnt += 1;
            if ((cilium_bend(C, x) > C.max_bend) && (C.midpnt > start)) {
                C.Tstroke = g_end;
                C.midpnt = start;
            }
        }
        if (C.midpnt > start)
            C.midpnt += tstep / C.m_speed;
        cil[c].midpnt = C.midpnt;
        cil[c].Tstroke = C.Tstroke;
    }
}

// The MD half of one MPCD step for one swimmer, in one workgroup: for each
// of the n + 1 force evaluations m, the beat's rest lengths are refreshed
// when the host's schedule says so, every term is evaluated in parallel
// into scratch, and each bead sums its terms in order, kicks v += F c_m
// (forces are accelerations, Matrix::add_mul3) and, for m < n, drifts
// x += v ststep into the other position buffer (MPCD::move, with the
// bounce-back at planar walls when the solvent has them).  Then the
// beat state machine.  The positions end in buffer cur ^ (n & 1).
__device__ __forceinline__ void md_step_block(const MdArgs &M, double *__restrict__ x0, double *__restrict__ x1,
                                              const MdSchedule &sched, int n, double dt, double g_end, double tstep,
                                              int walls, double ztop, double vwx, double vwy)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6, nwave = nt >> 6;
    __shared__ double s_hv[kMdBlock / 64][2][64 * 3];  // a wave's two chunks of a heavy bead's terms
    double *x = x0, *xn = x1;
    for (int m = 0; m <= n; ++m) {
        if ((sched.w[m >> 5] >> (m & 31)) & 1u) {
            md_phase_refresh(tid, nt, M.ncil, M.cil, x, M.newlen, M.oldlen);
            __syncthreads();
        }
        md_phase_terms(tid, nt, M.nterms, M.terms, M.tself, x, M.cil, M.newlen, M.contrib);
        __syncthreads();
        const MdKick K{(m == 0 || m == n) ? dt / 2.0 : dt, dt, ztop, vwx, vwy, m < n, walls};
        md_phase_light(tid, nt, M.nb, M.tstart, M.contrib, x, xn, M.v, M.f, M.mobile, K);
        for (int h = wave; h < M.nheavy; h += nwave)
            md_heavy_bead(M.heavy[h], lane, s_hv[wave], M.tstart, M.contrib, x, xn, M.v, M.f, M.mobile, K);
        __syncthreads();
        if (m < n) {
            double *t = x;
            x = xn;
            xn = t;
        }
    }
    md_phase_switch(tid, nt, M.ncil, M.cil, x, g_end, tstep);
}

__global__ __launch_bounds__(kMdBlock) void k_md_step(MdArgs M, double *__restrict__ x0, double *__restrict__ x1,
                                                  MdSchedule sched, int n, double dt, double g_end, double tstep,
                                                  int walls, double ztop, double vwx, double vwy)
{
    md_step_block(M, x0, x1, sched, n, dt, g_end, tstep, walls, ztop, vwx, vwy);
}

// Every swimmer of an ensemble, a workgroup each: the refresh schedule from
// the swimmer's own time (as md_run derives it on the host), the step, then
// the swimmer's position buffer and time advanced on the device.
__global__ __launch_bounds__(kMdBlock) void k_md_multi(MdBox *__restrict__ boxes, double tstep)
{
    MdBox &B = boxes[blockIdx.x];
    __shared__ MdSchedule s_sched;
    __shared__ double s_gend;
    if (threadIdx.x == 0)
        s_gend = md_schedule(B.globalt, B.dt, B.n, B.M.ncil > 0, s_sched);
    __syncthreads();
    const int cur = B.cur;
    md_step_block(B.M, B.x[cur], B.x[1 - cur], s_sched, B.n, B.dt, s_gend, tstep, B.walls, B.ztop, B.vw[0], B.vw[1]);
    __syncthreads();
    if (threadIdx.x == 0) {
        B.cur = cur ^ (B.n & 1);
        B.globalt = s_gend;
    }
}

// The wide path for large swimmers: one launch per phase over the whole GPU.

__global__ __launch_bounds__(kMdWideBlock) void k_md_refresh(MdArgs M, const double *__restrict__ x)
{
    md_phase_refresh(blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, M.ncil, M.cil, x, M.newlen,
                     M.oldlen);
}

__global__ __launch_bounds__(kMdWideBlock) void k_md_terms(MdArgs M, const double *__restrict__ x)
{
    md_phase_terms(blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, M.nterms, M.terms, M.tself, x,
                   M.cil, M.newlen, M.contrib);
}

// light beads on blocks [0, nlight_blocks), heavy beads one wave each after
__global__ __launch_bounds__(kMdWideBlock) void k_md_beads(MdArgs M, const double *__restrict__ x,
                                                       double *__restrict__ xn, MdKick K, int nlight_blocks)
{
    __shared__ double s_hv[kMdWideBlock / 64][2][64 * 3];
    if ((int)blockIdx.x < nlight_blocks) {
        md_phase_light(blockIdx.x * blockDim.x + threadIdx.x, nlight_blocks * blockDim.x, M.nb, M.tstart, M.contrib,
                       x, xn, M.v, M.f, M.mobile, K);
        return;
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int h = ((int)blockIdx.x - nlight_blocks) * (kMdWideBlock / 64) + wave;
    if (h < M.nheavy)
        md_heavy_bead(M.heavy[h], lane, s_hv[wave], M.tstart, M.contrib, x, xn, M.v, M.f, M.mobile, K);
}

__global__ __launch_bounds__(kMdWideBlock) void k_md_switch(MdArgs M, const double *__restrict__ x, double g_end,
                                                        double tstep)
{
    md_phase_switch(blockIdx.x * blockDim.x + threadIdx.x, gridDim.x * blockDim.x, M.ncil, M.cil, x, g_end, tstep);
}

// ---- the wide path over every swimmer of an ensemble: the workgroups of a
//      phase are split among the swimmers by prefix sums; a workgroup finds
//      its swimmer by binary search and takes that swimmer's items with a
//      stride of the swimmer's workgroups.  The position buffers at force
//      evaluation m are x[cur ^ (m & 1)] (n drifts flip them n times).

// the last swimmer whose first workgroup is at or before this one
__device__ __forceinline__ int md_swimmer_of_block(const int *__restrict__ blk, int nsw)
{
    int lo = 0, hi = nsw - 1;
    const int b = (int)blockIdx.x;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (blk[mid] <= b)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

__global__ void k_mdw_sched(MdBox *__restrict__ boxes, int nsw)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nsw) {
        MdBox &B = boxes[i];
        B.g_end = md_schedule(B.globalt, B.dt, B.n, B.M.ncil > 0, B.sched);
    }
}

__global__ __launch_bounds__(kMdWideBlock) void k_mdw_refresh(MdBox *__restrict__ boxes, const int *__restrict__ blk,
                                                           int nsw, int m)
{
    const int b = md_swimmer_of_block(blk, nsw);
    const MdBox &B = boxes[b];
    if (m > B.n || !((B.sched.w[m >> 5] >> (m & 31)) & 1u))
        return;
    const int i0 = ((int)blockIdx.x - blk[b]) * kMdWideBlock + threadIdx.x;
    md_phase_refresh(i0, (blk[b + 1] - blk[b]) * kMdWideBlock, B.M.ncil, B.M.cil, B.x[B.cur ^ (m & 1)], B.M.newlen,
                     B.M.oldlen);
}

__global__ __launch_bounds__(kMdWideBlock) void k_mdw_terms(MdBox *__restrict__ boxes, const int *__restrict__ blk,
                                                         int nsw, int m)
{
    const int b = md_swimmer_of_block(blk, nsw);
    const MdBox &B = boxes[b];
    if (m > B.n)
        return;
    const int i0 = ((int)blockIdx.x - blk[b]) * kMdWideBlock + threadIdx.x;
    md_phase_terms(i0, (blk[b + 1] - blk[b]) * kMdWideBlock, B.M.nterms, B.M.terms, B.M.tself, B.x[B.cur ^ (m & 1)],
                   B.M.cil, B.M.newlen, B.M.contrib);
}

__global__ __launch_bounds__(kMdWideBlock) void k_mdw_beads(MdBox *__restrict__ boxes, const int *__restrict__ blk,
                                                         const int *__restrict__ light, int nsw, int m)
{
    __shared__ double s_hv[kMdWideBlock / 64][2][64 * 3];
    const int b = md_swimmer_of_block(blk, nsw);
    const MdBox &B = boxes[b];
    if (m > B.n)
        return;
    const int n = B.n;
    const double dt = B.dt;
    const MdKick K{(m == 0 || m == n) ? dt / 2.0 : dt, dt, B.ztop, B.vw[0], B.vw[1], m < n, B.walls};
    const double *x = B.x[B.cur ^ (m & 1)];
    double *xn = B.x[B.cur ^ (m & 1) ^ 1];
    const int local = (int)blockIdx.x - blk[b];
    if (local < light[b]) {
        md_phase_light(local * kMdWideBlock + threadIdx.x, light[b] * kMdWideBlock, B.M.nb, B.M.tstart, B.M.contrib, x,
                       xn, B.M.v, B.M.f, B.M.mobile, K);
        return;
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int h = (local - light[b]) * (kMdWideBlock / 64) + wave;
    if (h < B.M.nheavy)
        md_heavy_bead(B.M.heavy[h], lane, s_hv[wave], B.M.tstart, B.M.contrib, x, xn, B.M.v, B.M.f, B.M.mobile, K);
}

__global__ __launch_bounds__(kMdWideBlock) void k_mdw_switch(MdBox *__restrict__ boxes, const int *__restrict__ blk,
                                                          int nsw, double tstep)
{
    const int b = md_swimmer_of_block(blk, nsw);
    const MdBox &B = boxes[b];
    const int i0 = ((int)blockIdx.x - blk[b]) * kMdWideBlock + threadIdx.x;
    md_phase_switch(i0, (blk[b + 1] - blk[b]) * kMdWideBlock, B.M.ncil, B.M.cil, B.x[B.cur ^ (B.n & 1)], B.g_end,
                    tstep);
}

// after every phase: the swimmer's buffer parity and time
__global__ void k_mdw_finish(MdBox *__restrict__ boxes, int nsw)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nsw) {
        MdBox &B = boxes[i];
        B.cur ^= B.n & 1;
        B.globalt = B.g_end;
    }
}

template <typename T>
hipError_t upload(T **d, const T *h, size_t n, hipStream_t st)
{
    hipError_t e = hipMalloc((void **)d, sizeof(T) * (n ? n : 1));
    if (e != hipSuccess)
        return e;
    if (n && h)
        return hipMemcpyAsync(*d, h, sizeof(T) * n, hipMemcpyHostToDevice, st);
    return hipSuccess;
}

}  // namespace

hipError_t md_alloc_upload(MdSystem &s, const double *pos3, const double *vel4, const unsigned char *mobile,
                           const std::vector<int> &tstart, const std::vector<MdTerm> &terms,
                           const std::vector<MdCilium> &cil, const std::vector<double> &newlen,
                           const std::vector<double> &oldlen, hipStream_t st)
{
    hipError_t e;
    const size_t nb = (size_t)s.nb;
    std::vector<int> tself(terms.size()), heavy;
    for (size_t b = 0; b < nb; ++b) {
        for (int t = tstart[b]; t < tstart[b + 1]; ++t)
            tself[(size_t)t] = (int)b;
        if (tstart[b + 1] - tstart[b] > kMdHeavy)
            heavy.push_back((int)b);
    }
    s.nheavy = (int)heavy.size();
    if ((e = upload(&s.x[0], pos3, nb * 3, st)) != hipSuccess || (e = upload(&s.x[1], pos3, nb * 3, st)) != hipSuccess ||
        (e = upload(&s.v, vel4, nb * 4, st)) != hipSuccess ||
        (e = upload(&s.f, (const double *)nullptr, nb * 3, st)) != hipSuccess ||
        (e = upload(&s.tstart, tstart.data(), tstart.size(), st)) != hipSuccess ||
        (e = upload(&s.terms, terms.data(), terms.size(), st)) != hipSuccess ||
        (e = upload(&s.tself, tself.data(), tself.size(), st)) != hipSuccess ||
        (e = upload(&s.contrib, (const double *)nullptr, terms.size() * 3, st)) != hipSuccess ||
        (e = upload(&s.cil, cil.data(), cil.size(), st)) != hipSuccess ||
        (e = upload(&s.newlen, newlen.data(), newlen.size(), st)) != hipSuccess ||
        (e = upload(&s.oldlen, oldlen.data(), oldlen.size(), st)) != hipSuccess ||
        (e = upload(&s.heavy, heavy.data(), heavy.size(), st)) != hipSuccess)
        return e;
    if (mobile && (e = upload(&s.mobile, mobile, nb, st)) != hipSuccess)
        return e;
    if ((e = hipMemsetAsync(s.f, 0, sizeof(double) * 3 * (nb ? nb : 1), st)) != hipSuccess)
        return e;
    // pageable sources: the copies must finish before the vectors go
    if ((e = hipStreamSynchronize(st)) != hipSuccess)
        return e;
    s.cur = 0;
    return hipSuccess;
}

MdArgs md_args(const MdSystem &s)
{
    return MdArgs{s.nb, s.nterms, s.ncil, s.nheavy, s.v, s.f, s.mobile, s.tstart, s.terms, s.tself,
                  s.cil, s.newlen, s.oldlen, s.contrib, s.heavy};
}

void md_advance_host(MdSystem &s, double tstep)
{
    const int n = md_substeps(s.ststep, tstep);
    MdSchedule sched;
    s.globalt = md_schedule(s.globalt, s.ststep, n, s.ncil > 0, sched);
    if (n & 1)
        s.cur = 1 - s.cur;
}

void md_launch_multi(MdBox *boxes, int nsw, double tstep, hipStream_t st)
{
    if (nsw > 0)
        hipLaunchKernelGGL(k_md_multi, dim3(nsw), dim3(kMdBlock), 0, st, boxes, tstep);
}

void md_launch_multi_wide(MdBox *boxes, const MdWide &w, double tstep, hipStream_t st)
{
    if (w.nsw <= 0)
        return;
    const dim3 small((w.nsw + 63) / 64), blk(kMdWideBlock);
    hipLaunchKernelGGL(k_mdw_sched, small, dim3(64), 0, st, boxes, w.nsw);
    for (int m = 0; m <= w.nmax; ++m) {
        if (w.cil_total)
            hipLaunchKernelGGL(k_mdw_refresh, dim3(w.cil_total), blk, 0, st, boxes, w.cil_blk, w.nsw, m);
        if (w.term_total)
            hipLaunchKernelGGL(k_mdw_terms, dim3(w.term_total), blk, 0, st, boxes, w.term_blk, w.nsw, m);
        hipLaunchKernelGGL(k_mdw_beads, dim3(w.bead_total), blk, 0, st, boxes, w.bead_blk, w.light_blk, w.nsw, m);
    }
    if (w.cil_total)
        hipLaunchKernelGGL(k_mdw_switch, dim3(w.cil_total), blk, 0, st, boxes, w.cil_blk, w.nsw, tstep);
    hipLaunchKernelGGL(k_mdw_finish, small, dim3(64), 0, st, boxes, w.nsw);
}

void md_free(MdSystem &s)
{
    void *p[] = {s.x[0], s.x[1], s.v,      s.f,      s.mobile, s.tstart, s.terms,
                 s.tself, s.contrib, s.cil, s.newlen, s.oldlen, s.heavy};
    for (void *q : p)
        if (q)
            (void)hipFree(q);
    s = MdSystem{};
}

hipError_t md_run(MdSystem &s, double tstep, hipStream_t st)
{
    if (s.nb == 0)
        return hipSuccess;
    const int n = md_substeps(s.ststep, tstep);
    if (n < 1 || n >= kMdMaxSubsteps)
        return hipErrorInvalidValue;
    const double dt = s.ststep;
    MdSchedule sched;
    const double g = md_schedule(s.globalt, dt, n, s.ncil > 0, sched);
    MdArgs M{s.nb, s.nterms, s.ncil, s.nheavy, s.v, s.f, s.mobile, s.tstart, s.terms, s.tself,
             s.cil, s.newlen, s.oldlen, s.contrib, s.heavy};
    if (s.nterms <= kMdWideTerms) {
        hipLaunchKernelGGL(k_md_step, dim3(1), dim3(kMdBlock), 0, st, M, s.x[s.cur], s.x[1 - s.cur], sched, n, dt, g,
                           tstep, s.walls, s.ztop, s.vw[0], s.vw[1]);
        if (n & 1)  // n drifts swap the buffers n times
            s.cur = 1 - s.cur;
    } else {
        // large swimmers: each phase of each evaluation over the whole GPU
        const auto blocks = [](long long items) { return dim3((unsigned)((items + kMdWideBlock - 1) / kMdWideBlock)); };
        const int nlight = (int)blocks(s.nb).x;
        const int nheavy_blocks = (s.nheavy + kMdWideBlock / 64 - 1) / (kMdWideBlock / 64);
        for (int m = 0; m <= n; ++m) {
            double *x = s.x[s.cur], *xn = s.x[1 - s.cur];
            if ((sched.w[m >> 5] >> (m & 31)) & 1u)
                hipLaunchKernelGGL(k_md_refresh, blocks(s.ncil), dim3(kMdWideBlock), 0, st, M, x);
            hipLaunchKernelGGL(k_md_terms, blocks(s.nterms), dim3(kMdWideBlock), 0, st, M, x);
            const MdKick K{(m == 0 || m == n) ? dt / 2.0 : dt, dt, s.ztop, s.vw[0], s.vw[1], m < n, s.walls};
            hipLaunchKernelGGL(k_md_beads, dim3(nlight + nheavy_blocks), dim3(kMdWideBlock), 0, st, M, x, xn, K,
                               nlight);
            if (m < n)
                s.cur = 1 - s.cur;
        }
        hipLaunchKernelGGL(k_md_switch, blocks(s.ncil), dim3(kMdWideBlock), 0, st, M, s.x[s.cur], g, tstep);
    }
    s.globalt = g;
    return hipGetLastError();
}

}  // namespace mcm
