// Bead-spring MD sub-steps on gfx950 (see mpcd_md.hpp).
//
// All arithmetic is FP64 and, compiled with -ffp-contract=off, rounds each
// operation separately in the reference's order.  pow(x, 2.0) and
// pow(x, 2.5) of the beat rule are x*x and (x*x)*sqrt(x), what a
// -ffast-math build of the reference computes.
#include <algorithm>
#include <cmath>

#include "mpcd_device.hpp"
#include "mpcd_md.hpp"

#include <cstring>

namespace mcm {

namespace {

constexpr int kMdBlock = 1024;
constexpr int kMdHeavy = 32;     // beads with more terms are summed by a wave or a workgroup
constexpr int kMdWaveTerms = 256;  // beads with more terms are summed by a workgroup (the sphere's centre)
constexpr int kMdRefLanes = 8;  // threads per cilium in the rest-length refresh

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
// bead's running sum rounds exactly as the reference's += / -= do.  xs is
// the bead's own position; a cilium term's rest length reads newlen[T.idx]
// with the cilium's bond length in T.b.
__device__ __forceinline__ void md_term(const MdTerm &T, const double (&xs)[3], const double *__restrict__ x,
                                        const double *__restrict__ newlen, const double *__restrict__ ktab,
                                        double (&out)[3])
{
    const double *xq = x + 3 * (long long)T.partner;
    const double xp[3] = {xq[0], xq[1], xq[2]};
    const int flags = term_flags(T);
    const double k = ktab[term_kidx(T)];
    double rest = T.b;
    const int kind = flags & kTermKindMask;
    if (kind != kTermConst) {
        const double nl = newlen[term_idx(T)];
        rest = kind == kTermNewLen ? nl : sqrt(T.b * T.b + T.b * nl);
    }
    if (flags & kTermAttach) {
        const bool second = (flags & kTermAttachB) != 0;
        double p1[3], p0[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            p1[i] = second ? xp[i] : xs[i];
            p0[i] = second ? xs[i] : xp[i];
        }
        const double d = sqrt(md_sqdist(p1, p0));
        double F[3] = {0.0, 0.0, 0.0};
        if (d != 0) {
            const double fac = -2.0 * k * (1.0 - rest / d);
#pragma unroll
            for (int i = 0; i < 3; ++i)
                F[i] = (p1[i] - p0[i]) * fac;
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
            out[i] = second ? -F[i] : F[i];
    } else {
        const double fac = -2.0 * k * (1.0 - rest / sqrt(md_sqdist(xs, xp)));
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

// A bead's own state, read before its terms: position, velocity and
// whether it moves.
__device__ __forceinline__ void md_bead_load(int b, const double *__restrict__ x, const double *__restrict__ v,
                                             const unsigned char *__restrict__ mobile, double (&xs)[3],
                                             double (&v0)[3], bool &mob)
{
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        xs[i] = x[3 * (long long)b + i];
        v0[i] = v[4 * (long long)b + i];
    }
    mob = !mobile || mobile[b];
}

// Kick (v += F c) and, for sub-steps, drift into the other buffer (with the
// walls' bounce-back); fixed beads stay put.  The bead's force is stored.
// xs, v0, mob: the bead's state from md_bead_load.
__device__ __forceinline__ void md_kick_drift(int b, const double f[3], const double (&xs)[3],
                                              const double (&v0)[3], bool mob, double c, bool drift, double dt,
                                              double *__restrict__ xn, double *__restrict__ v,
                                              double *__restrict__ fo, int walls, double ztop, double vwx,
                                              double vwy)
{
    double *vb = v + 4 * (long long)b;
    double *xb = xn + 3 * (long long)b;
#pragma unroll
    for (int i = 0; i < 3; ++i)
        fo[3 * (long long)b + i] = f[i];
    if (mob) {
        double nv[3], nx[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
            nv[i] = v0[i] + f[i] * c;
        if (drift) {
#pragma unroll
            for (int i = 0; i < 3; ++i)
                nx[i] = xs[i] + nv[i] * dt;
            if (walls)
                wall_bounce(walls, ztop, vwx, vwy, nx[0], nx[1], nx[2], nv[0], nv[1], nv[2]);
#pragma unroll
            for (int i = 0; i < 3; ++i)
                xb[i] = nx[i];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i)
            vb[i] = nv[i];
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

struct MdKick {
    double c, dt, ztop, vwx, vwy;
    int drift, walls;
};

// A swimmer's arrays for one force evaluation.
struct MdEval {
    const int *__restrict__ tstart;
    const MdTerm *__restrict__ terms;
    const double *__restrict__ ktab;
    const MdCilium *__restrict__ cil;
    const double *__restrict__ newlen;
    const double *__restrict__ x;
    double *__restrict__ xn;
    double *__restrict__ v;
    double *__restrict__ f;
    const unsigned char *__restrict__ mobile;
};

__device__ __forceinline__ void md_wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

constexpr int kMdLanes = 4;  // lanes per light bead

// Light bead b (at most kMdHeavy terms; on = false: none), kMdLanes lanes:
// lane `sub` of the group evaluates terms sub, sub + kMdLanes, ... of the
// bead's list, a batch of kMdLanes at a time, into the wave's LDS chunk hv
// (each batch's term records are read while the previous batch is
// evaluated); the group's lanes 0, 1, 2 add the batch's x, y, z components
// in list order (the value each term adds is the one the reference's loop
// adds), then lane 0 kicks and drifts.  Every lane of the wave calls this
// together (the batches are wave-uniform).
__device__ __forceinline__ void md_light_bead(int b, bool on, int lane, double *hv, const MdEval &E,
                                              const MdKick &K)
{
    const int sub = lane & (kMdLanes - 1), base = lane - sub;
    int t0 = 0, n = 0;
    if (on) {
        t0 = E.tstart[b];
        n = E.tstart[b + 1] - t0;
        if (n > kMdHeavy) {  // summed by a workgroup
            n = 0;
            on = false;
        }
    }
    double xs[3] = {0.0, 0.0, 0.0}, v0[3] = {0.0, 0.0, 0.0};
    bool mob = false;
    if (on)
        md_bead_load(b, E.x, E.v, E.mobile, xs, v0, mob);
    const MdTerm *tl = E.terms + t0;
    MdTerm tn{};
    if (sub < n)
        tn = tl[sub];
    double acc = 0.0;  // component `sub` of the force (sub < 3)
    for (int c0 = 0; __any(c0 < n); c0 += kMdLanes) {
        const MdTerm T = tn;
        if (c0 + kMdLanes + sub < n)
            tn = tl[c0 + kMdLanes + sub];
        if (c0 + sub < n) {
            double c[3];
            md_term(T, xs, E.x, E.newlen, E.ktab, c);
            hv[3 * lane] = c[0];
            hv[3 * lane + 1] = c[1];
            hv[3 * lane + 2] = c[2];
        }
        md_wave_sync();
        if (sub < 3) {
            const int m = min(kMdLanes, n - c0);
            for (int l = 0; l < m; ++l)
                acc += hv[3 * (base + l) + sub];
        }
        md_wave_sync();
    }
    double f[3];
    f[0] = __shfl(acc, base);
    f[1] = __shfl(acc, base + 1);
    f[2] = __shfl(acc, base + 2);
    if (on && sub == 0)
        md_kick_drift(b, f, xs, v0, mob, K.c, K.drift, K.dt, E.xn, E.v, E.f, K.walls, K.ztop, K.vwx, K.vwy);
}

// A bead with a long term list (at most kMdWaveTerms), by one wave (every
// lane calls this): 64 terms at a time evaluated into the wave's LDS chunk
// hv (the next batch's term records read meanwhile); lanes 0, 1, 2 add the
// x, y, z components in list order, then lane 0 kicks and drifts.
__device__ __forceinline__ void md_wave_bead(int b, int lane, double *hv, const MdEval &E, const MdKick &K)
{
    const int t0 = E.tstart[b], n = E.tstart[b + 1] - t0;
    double xs[3], v0[3];
    bool mob;
    md_bead_load(b, E.x, E.v, E.mobile, xs, v0, mob);
    const MdTerm *tl = E.terms + t0;
    MdTerm tn{};
    if (lane < n)
        tn = tl[lane];
    double acc = 0.0;  // component `lane` of the force (lanes 0..2)
    for (int c0 = 0; c0 < n; c0 += 64) {
        const MdTerm T = tn;
        if (c0 + 64 + lane < n)
            tn = tl[c0 + 64 + lane];
        if (c0 + lane < n) {
            double c[3];
            md_term(T, xs, E.x, E.newlen, E.ktab, c);
            hv[3 * lane] = c[0];
            hv[3 * lane + 1] = c[1];
            hv[3 * lane + 2] = c[2];
        }
        md_wave_sync();
        if (lane < 3) {
            const int m = min(64, n - c0);
            int j = 0;
            for (; j + 8 <= m; j += 8) {
                double cv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    cv[u] = hv[3 * (j + u) + lane];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    acc += cv[u];
            }
            for (; j < m; ++j)
                acc += hv[3 * j + lane];
        }
        md_wave_sync();
    }
    double f[3];
    f[0] = __shfl(acc, 0);
    f[1] = __shfl(acc, 1);
    f[2] = __shfl(acc, 2);
    if (lane == 0)
        md_kick_drift(b, f, xs, v0, mob, K.c, K.drift, K.dt, E.xn, E.v, E.f, K.walls, K.ztop, K.vwx, K.vwy);
}

constexpr int kMdHeavyChunk = 1024;  // a heavy bead's terms evaluated together into LDS

// A bead with a long term list, by a whole workgroup (every thread calls
// this): up to kMdHeavyChunk terms at a time evaluated in parallel into
// LDS (hv), two per thread with their records read together; lanes 0, 1, 2
// of the first wave add the x, y, z components in list order (one
// instruction advances all three sums), then thread 0 kicks and drifts.
__device__ __forceinline__ void md_heavy_bead(int b, double *hv, const MdEval &E, const MdKick &K)
{
    const int t0 = E.tstart[b], t1 = E.tstart[b + 1];
    const int q = threadIdx.x, nt = blockDim.x;
    double xs[3], v0[3];
    bool mob;
    md_bead_load(b, E.x, E.v, E.mobile, xs, v0, mob);
    double acc = 0.0;  // component q of the force (threads 0..2)
    for (int cb = t0; cb < t1; cb += kMdHeavyChunk) {
        const int m = min(kMdHeavyChunk, t1 - cb);
        for (int j = q; j < m; j += 2 * nt) {
            const bool two = j + nt < m;
            const MdTerm A = E.terms[cb + j];
            MdTerm B{};
            if (two)
                B = E.terms[cb + j + nt];
            double c[3];
            md_term(A, xs, E.x, E.newlen, E.ktab, c);
            hv[3 * j] = c[0];
            hv[3 * j + 1] = c[1];
            hv[3 * j + 2] = c[2];
            if (two) {
                md_term(B, xs, E.x, E.newlen, E.ktab, c);
                hv[3 * (j + nt)] = c[0];
                hv[3 * (j + nt) + 1] = c[1];
                hv[3 * (j + nt) + 2] = c[2];
            }
        }
        __syncthreads();
        if (q < 3) {
            int j = 0;
            for (; j + 8 <= m; j += 8) {
                double cv[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    cv[u] = hv[3 * (j + u) + q];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    acc += cv[u];
            }
            for (; j < m; ++j)
                acc += hv[3 * j + q];
        }
        __syncthreads();
    }
    double f[3];
    f[0] = __shfl(acc, 0);
    f[1] = __shfl(acc, 1);
    f[2] = __shfl(acc, 2);
    if (threadIdx.x == 0)
        md_kick_drift(b, f, xs, v0, mob, K.c, K.drift, K.dt, E.xn, E.v, E.f, K.walls, K.ztop, K.vwx, K.vwy);
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
// when the schedule says so, then each bead evaluates its terms in
// parallel (8 lanes, or the workgroup for a long list) and adds them in list
// order, kicks v += F c_m (forces are accelerations, Matrix::add_mul3) and,
// for m < n, drifts x += v ststep into the other position buffer
// (MPCD::move, with the bounce-back at planar walls when the solvent has
// them); terms read only the current buffer.  Then the beat state machine.
// The positions end in buffer cur ^ (n & 1).
__device__ __forceinline__ void md_step_block(const MdArgs &M, double *__restrict__ x0, double *__restrict__ x1,
                                              const MdSchedule &sched, int n, double dt, double g_end, double tstep,
                                              int walls, double ztop, double vwx, double vwy)
{
    const int tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6;
    // light beads: a wave's batch of 64 term values; heavy beads: a chunk
    __shared__ double s_hv[(kMdBlock / 64) * 64 * 3 > kMdHeavyChunk * 3 ? (kMdBlock / 64) * 64 * 3 : kMdHeavyChunk * 3];
    double *x = x0, *xn = x1;
    for (int m = 0; m <= n; ++m) {
        if ((sched.w[m >> 5] >> (m & 31)) & 1u) {
            md_phase_refresh(tid, nt, M.ncil, M.cil, x, M.newlen, M.oldlen);
            __syncthreads();
        }
        const MdKick K{(m == 0 || m == n) ? dt / 2.0 : dt, dt, ztop, vwx, vwy, m < n, walls};
        const MdEval E{M.tstart, M.terms, M.ktab, M.cil, M.newlen, x, xn, M.v, M.f, M.mobile};
        for (int r = 0; r * (nt / kMdLanes) < M.nb; ++r) {
            const int b = r * (nt / kMdLanes) + tid / kMdLanes;
            md_light_bead(b, b < M.nb, lane, s_hv + wave * 64 * 3, E, K);
        }
        __syncthreads();
        for (int h = 0; h < M.nheavy; ++h)
            md_heavy_bead(M.heavy[h], s_hv, E, K);
        __syncthreads();
        if (m < n) {
            double *t = x;
            x = xn;
            xn = t;
        }
    }
    md_phase_switch(tid, nt, M.ncil, M.cil, x, g_end, tstep);
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

// ---- the wide path, for one swimmer or every swimmer of an ensemble: each
//      phase of each force evaluation is one launch; its workgroups are
//      split among the swimmers by prefix sums; a workgroup finds its
//      swimmer by binary search and takes that swimmer's items with a
//      stride of the swimmer's workgroups.  The position buffers at force
//      evaluation m are x[cur ^ (m & 1)] (n drifts flip them n times).  The
//      schedule, buffer parity and time live in the swimmer's MdBox, so a
//      captured graph replays whole MD halves.

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

// The refresh with kMdRefLanes threads per cilium, one rod-0 segment each
// (a segment's new rest length reads its own old one only).  BLK: threads
// per workgroup of the wide launches.
template <int BLK>
__global__ __launch_bounds__(BLK) void k_mdw_refresh(MdBox *__restrict__ boxes, const int *__restrict__ blk, int nsw,
                                                   int m)
{
    const int b = md_swimmer_of_block(blk, nsw);
    const MdBox &B = boxes[b];
    if (m > B.n || !((B.sched.w[m >> 5] >> (m & 31)) & 1u))
        return;
    const int i0 = ((int)blockIdx.x - blk[b]) * BLK + threadIdx.x;
    const int stride = (blk[b + 1] - blk[b]) * BLK;
    const double *x = B.x[B.cur ^ (m & 1)];
    for (int it = i0; it < B.M.ncil * kMdRefLanes; it += stride) {
        const MdCilium &C = B.M.cil[it / kMdRefLanes];
        for (int i = C.num_anker - 1 + it % kMdRefLanes; i < C.rodlength - 1; i += kMdRefLanes) {
            const double l = cilium_length(C, x, B.M.oldlen, i);
            B.M.newlen[C.len_off + i] = l;
            B.M.oldlen[C.len_off + i] = l;
        }
    }
}

template <int BLK>
__global__ __launch_bounds__(BLK) void k_mdw_beads(MdBox *__restrict__ boxes, const int *__restrict__ blk, int nsw,
                                                 int m)
{
    constexpr int kWaves = BLK / 64;
    // light beads: a wave's batch of 64 term values; heavy beads: a chunk
    __shared__ double s_hv[kMdHeavyChunk * 3];
    static_assert(kMdHeavyChunk * 3 >= (BLK / 64) * 64 * 3, "light staging fits");
    const int b = md_swimmer_of_block(blk, nsw);
    const MdBox &B = boxes[b];
    if (m > B.n)
        return;
    const int n = B.n;
    const double dt = B.dt;
    const MdKick K{(m == 0 || m == n) ? dt / 2.0 : dt, dt, B.ztop, B.vw[0], B.vw[1], m < n, B.walls};
    const double *x = B.x[B.cur ^ (m & 1)];
    double *xn = B.x[B.cur ^ (m & 1) ^ 1];
    const MdEval E{B.M.tstart, B.M.terms, B.M.ktab, B.M.cil, B.M.newlen, x, xn, B.M.v, B.M.f, B.M.mobile};
    // the swimmer's beads with the longest term lists first, a workgroup
    // each; then the other long lists, a wave each; then the light beads,
    // BLK / kMdLanes per workgroup
    const int local = (int)blockIdx.x - blk[b];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int big = B.M.nbig, nwb = (B.M.nheavy - big + kWaves - 1) / kWaves;
    if (local < big) {
        md_heavy_bead(B.M.heavy[local], s_hv, E, K);
        return;
    }
    if (local < big + nwb) {
        const int h = big + (local - big) * kWaves + wave;
        if (h < B.M.nheavy)
            md_wave_bead(B.M.heavy[h], lane, s_hv + wave * 64 * 3, E, K);
        return;
    }
    const int bead = (local - big - nwb) * (BLK / kMdLanes) + (int)threadIdx.x / kMdLanes;
    md_light_bead(bead, bead < B.M.nb, lane, s_hv + wave * 64 * 3, E, K);
}

template <int BLK>
__global__ __launch_bounds__(BLK) void k_mdw_switch(MdBox *__restrict__ boxes, const int *__restrict__ blk, int nsw,
                                                  double tstep)
{
    const int b = md_swimmer_of_block(blk, nsw);
    const MdBox &B = boxes[b];
    const int i0 = ((int)blockIdx.x - blk[b]) * BLK + threadIdx.x;
    md_phase_switch(i0, (blk[b + 1] - blk[b]) * BLK, B.M.ncil, B.M.cil, B.x[B.cur ^ (B.n & 1)], B.g_end,
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
                           const std::vector<double> &ktab, const std::vector<MdCilium> &cil,
                           const std::vector<double> &newlen, const std::vector<double> &oldlen, hipStream_t st)
{
    hipError_t e;
    const size_t nb = (size_t)s.nb;
    // beads with long term lists, the longest first
    std::vector<int> heavy;
    for (size_t b = 0; b < nb; ++b) {
        if (tstart[b + 1] - tstart[b] > kMdHeavy)
            heavy.push_back((int)b);
    }
    const auto len = [&](int b) { return tstart[b + 1] - tstart[b]; };
    std::stable_sort(heavy.begin(), heavy.end(), [&](int a, int c) { return len(a) > len(c); });
    s.nheavy = (int)heavy.size();
    s.nbig = 0;
    while (s.nbig < s.nheavy && len(heavy[s.nbig]) > kMdWaveTerms)
        ++s.nbig;
    if ((e = upload(&s.x[0], pos3, nb * 3, st)) != hipSuccess || (e = upload(&s.x[1], pos3, nb * 3, st)) != hipSuccess ||
        (e = upload(&s.v, vel4, nb * 4, st)) != hipSuccess ||
        (e = upload(&s.f, (const double *)nullptr, nb * 3, st)) != hipSuccess ||
        (e = upload(&s.tstart, tstart.data(), tstart.size(), st)) != hipSuccess ||
        (e = upload(&s.terms, terms.data(), terms.size(), st)) != hipSuccess ||
        (e = upload(&s.ktab, ktab.data(), ktab.size(), st)) != hipSuccess ||
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
    return MdArgs{s.nb, s.nterms, s.ncil, s.nheavy, s.v, s.f, s.mobile, s.tstart, s.terms, s.ktab,
                  s.cil, s.newlen, s.oldlen, s.heavy, s.nbig};
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

template <int BLK>
static void md_launch_wide(MdBox *boxes, const MdWide &w, double tstep, hipStream_t st, const MdSchedule *only)
{
    const dim3 small((w.nsw + 63) / 64), blk(BLK);
    hipLaunchKernelGGL(k_mdw_sched, small, dim3(64), 0, st, boxes, w.nsw);
    for (int m = 0; m <= w.nmax; ++m) {
        const bool due = !only || ((only->w[m >> 5] >> (m & 31)) & 1u);
        if (w.cil_total && due)
            hipLaunchKernelGGL(k_mdw_refresh<BLK>, dim3(w.ref_total), blk, 0, st, boxes, w.ref_blk, w.nsw, m);
        hipLaunchKernelGGL(k_mdw_beads<BLK>, dim3(w.bead_total), blk, 0, st, boxes, w.bead_blk, w.nsw, m);
    }
    if (w.cil_total)
        hipLaunchKernelGGL(k_mdw_switch<BLK>, dim3(w.cil_total), blk, 0, st, boxes, w.cil_blk, w.nsw, tstep);
    hipLaunchKernelGGL(k_mdw_finish, small, dim3(64), 0, st, boxes, w.nsw);
}

void md_launch_multi_wide(MdBox *boxes, const MdWide &w, double tstep, hipStream_t st, const MdSchedule *only)
{
    if (w.nsw <= 0)
        return;
    switch (w.block) {  // the tables were made for this workgroup size
    case 256: md_launch_wide<256>(boxes, w, tstep, st, only); break;
    case 1024: md_launch_wide<1024>(boxes, w, tstep, st, only); break;
    default: md_launch_wide<512>(boxes, w, tstep, st, only); break;
    }
}

void md_free(MdSystem &s)
{
    s.graphs.clear();
    void *p[] = {s.x[0], s.x[1], s.v,      s.f,     s.mobile, s.tstart, s.terms, s.ktab,
                 s.cil,  s.newlen, s.oldlen, s.heavy, s.dbox,   s.dtab};
    for (void *q : p)
        if (q)
            (void)hipFree(q);
    s = MdSystem{};
}

MdBox md_box(const MdSystem &s, double tstep)
{
    MdBox m{};
    m.M = md_args(s);
    m.x[0] = s.x[0];
    m.x[1] = s.x[1];
    m.cur = s.cur;
    m.n = md_substeps(s.ststep, tstep);
    m.dt = s.ststep;
    m.globalt = s.globalt;
    m.walls = s.walls;
    m.ztop = s.ztop;
    m.vw[0] = s.vw[0];
    m.vw[1] = s.vw[1];
    return m;
}

void md_wide_tables(const MdBox *boxes, int nsw, std::vector<int> &tab, MdWide &w, int block)
{
    if (!md_block_ok(block))
        block = kMdWideBlock;
    const int waves = block / 64;
    tab.assign((size_t)(4 * nsw + 3), 0);
    int *cil = tab.data(), *bead = cil + nsw + 1, *light = bead + nsw + 1, *ref = light + nsw;
    const auto blocks = [block](long long items) { return (int)((items + block - 1) / block); };
    int nmax = 0;
    for (int i = 0; i < nsw; ++i) {
        const MdBox &m = boxes[i];
        cil[i + 1] = cil[i] + blocks(m.M.ncil);
        ref[i + 1] = ref[i] + blocks((long long)m.M.ncil * kMdRefLanes);
        light[i] = blocks((long long)m.M.nb * kMdLanes);
        bead[i + 1] = bead[i] + light[i] + m.M.nbig + (m.M.nheavy - m.M.nbig + waves - 1) / waves;
        nmax = std::max(nmax, m.n);
    }
    w = MdWide{nullptr, nullptr, nullptr, cil[nsw], bead[nsw], nullptr, ref[nsw], nmax, nsw, block};
}

void md_wide_bind(MdWide &w, const int *dtab)
{
    const int nsw = w.nsw;
    w.cil_blk = dtab;
    w.bead_blk = dtab + (nsw + 1);
    w.light_blk = dtab + 2 * (nsw + 1);
    w.ref_blk = dtab + 3 * nsw + 2;
}

// The swimmer's device state from the host's and its tables for steps of
// length tstep; the graphs captured for another state or length go (the
// stream is idle).
static hipError_t md_prepare(MdSystem &s, double tstep)
{
    hipError_t e;
    s.graphs.clear();
    s.graph_tstep = -1.0;
    const MdBox m = md_box(s, tstep);
    std::vector<int> tab;
    md_wide_tables(&m, 1, tab, s.wide);
    if (!s.dbox && (e = hipMalloc((void **)&s.dbox, sizeof(MdBox))) != hipSuccess)
        return e;
    if (!s.dtab && (e = hipMalloc((void **)&s.dtab, sizeof(int) * 8)) != hipSuccess)
        return e;
    if ((e = hipMemcpy(s.dbox, &m, sizeof m, hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemcpy(s.dtab, tab.data(), sizeof(int) * tab.size(), hipMemcpyHostToDevice)) != hipSuccess)
        return e;
    md_wide_bind(s.wide, s.dtab);
    s.graph_tstep = tstep;
    return hipSuccess;
}

// The graph of one MD half of length tstep with the refreshes of sched.
static hipError_t md_capture(MdSystem &s, double tstep, const MdSchedule &sched, hipStream_t st, hipGraph_t &g,
                             hipGraphExec_t &x)
{
    hipError_t e;
    if ((e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal)) != hipSuccess)
        return e;
    md_launch_multi_wide(s.dbox, s.wide, tstep, st, &sched);
    const hipError_t le = hipGetLastError();
    e = hipStreamEndCapture(st, &g);
    if (le != hipSuccess || e != hipSuccess) {
        if (g)
            (void)hipGraphDestroy(g);
        g = nullptr;
        return le != hipSuccess ? le : e;
    }
    if ((e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0)) != hipSuccess) {
        (void)hipGraphDestroy(g);
        g = nullptr;
        x = nullptr;
    }
    return e;
}

hipError_t md_run(MdSystem &s, double tstep, hipStream_t st)
{
    if (s.nb == 0)
        return hipSuccess;
    const int n = md_substeps(s.ststep, tstep);
    if (n < 1 || n >= kMdMaxSubsteps)
        return hipErrorInvalidValue;
    hipError_t e;
    if (tstep != s.graph_tstep &&
        ((e = hipStreamSynchronize(st)) != hipSuccess || (e = md_prepare(s, tstep)) != hipSuccess))
        return e;
    // the refreshes of this step, from the host's globalt (the device's is
    // the same: both advance it by the same additions); each schedule has a
    // graph of its own (the refresh phase is a no-op launch elsewhere, which
    // the graph leaves out; its kernel still checks the device's schedule)
    MdSchedule sched;
    md_schedule(s.globalt, s.ststep, n, s.ncil > 0, sched);
    hipGraphExec_t x = nullptr;
    e = s.graphs.get(sched, n,
                     [&](const MdSchedule &use, hipGraph_t &g, hipGraphExec_t &ex) {
                         return md_capture(s, tstep, use, st, g, ex);
                     },
                     x);
    if (e != hipSuccess || (e = hipGraphLaunch(x, st)) != hipSuccess)
        return e;
    md_advance_host(s, tstep);
    return hipSuccess;
}

}  // namespace mcm
