// Bead-spring MD sub-steps on gfx950 (see mpcd_md.hpp).
//
// All arithmetic is FP64 and, compiled with -ffp-contract=off, rounds each
// operation separately in the reference's order.  pow(x, 2.0) and
// pow(x, 2.5) of the beat rule are x*x and (x*x)*sqrt(x), what a
// -ffast-math build of the reference computes.
#include <cmath>

#include "mpcd_md.hpp"

namespace mcm {

namespace {

constexpr int kMdBlock = 256;

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

// One bead: its force from its ordered terms (MDObject::spring,
// src/libcilia/mdobject.cpp:8-17; attachments as Clami::global_forces,
// src/clami.cpp:188-198), then v += F c (Matrix::add_mul3) and, when drift
// is set, x += v ststep (MPCD::move, src/libcilia/mpcd.cpp:340-343) into the
// other buffer.  Fixed beads keep their state.
__global__ __launch_bounds__(kMdBlock) void k_md_substep(int nb, const double *__restrict__ x, double *__restrict__ xo,
                                                         double *__restrict__ v, double *__restrict__ fo,
                                                         const unsigned char *__restrict__ mobile,
                                                         const int *__restrict__ tstart, const MdTerm *__restrict__ terms,
                                                         const MdCilium *__restrict__ cil,
                                                         const double *__restrict__ newlen, double c, double dt,
                                                         int drift)
{
    const int b = blockIdx.x * kMdBlock + threadIdx.x;
    if (b >= nb)
        return;
    const double *xs = x + 3 * (long long)b;
    double f[3] = {0.0, 0.0, 0.0};
    for (int t = tstart[b]; t < tstart[b + 1]; ++t) {
        const MdTerm T = terms[t];
        const double *xp = x + 3 * (long long)T.partner;
        double rest = T.b;
        const int kind = T.flags & kTermKindMask;
        if (kind != kTermConst) {
            const MdCilium &C = cil[T.cil];
            const double nl = newlen[C.len_off + T.idx];
            rest = kind == kTermNewLen ? nl : sqrt(C.bondlength * C.bondlength + C.bondlength * nl);
        }
        if (T.flags & kTermAttach) {
            // Vector3d spring(pos1 = sphere point, pos0 = cilium point)
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
                f[i] = second ? f[i] - F[i] : f[i] + F[i];
        } else {
            const double fac = -2.0 * T.k * (1.0 - rest / sqrt(md_sqdist(xs, xp)));
#pragma unroll
            for (int i = 0; i < 3; ++i)
                f[i] += (xs[i] - xp[i]) * fac;
        }
    }
    double *vb = v + 4 * (long long)b;
    double *xb = xo + 3 * (long long)b;
    fo[3 * (long long)b] = f[0];
    fo[3 * (long long)b + 1] = f[1];
    fo[3 * (long long)b + 2] = f[2];
    if (!mobile || mobile[b]) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
            vb[i] += f[i] * c;
        if (drift) {
#pragma unroll
            for (int i = 0; i < 3; ++i)
                xb[i] = xs[i] + vb[i] * dt;
        }
    } else if (drift) {
#pragma unroll
        for (int i = 0; i < 3; ++i)
            xb[i] = xs[i];
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

// The rest-length refresh at the top of Cilia::recalcforce (cilia.cpp:207-215).
__global__ void k_md_cilia_len(int ncil, const MdCilium *__restrict__ cil, const double *__restrict__ x,
                               double *__restrict__ newlen, double *__restrict__ oldlen)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncil)
        return;
    const MdCilium C = cil[c];
    for (int i = C.num_anker - 1; i < C.rodlength - 1; ++i) {
        const double l = cilium_length(C, x, oldlen, i);
        newlen[C.len_off + i] = l;
        oldlen[C.len_off + i] = l;
    }
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

// Beat state machine: Cilia::switch_logic (src/cilia.cpp:241-258) or
// ConstantBeatCilia::switch_logic (src/ConstantBeatCilia.cpp:8-23).
__global__ void k_md_switch(int ncil, MdCilium *__restrict__ cil, const double *__restrict__ x, double globalt,
                            double tstep)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncil)
        return;
    MdCilium C = cil[c];
    const double start = C.num_anker;
    if (C.constant_beat) {
        if (globalt - C.Tstroke >= C.power_stroke_duration && C.midpnt == start)
            C.midpnt += 1;
        if (globalt - C.Tstroke - C.power_stroke_duration >= C.recovery_stroke_duration && C.midpnt > start) {
            C.Tstroke = globalt;
            C.midpnt = start;
        }
    } else {
        if ((C.midpnt == start) && globalt - C.Tstroke > 100.0)
            C.midpnt += 1;
        if ((cilium_bend(C, x) > C.max_bend) && (C.midpnt > start)) {
            C.Tstroke = globalt;
            C.midpnt = start;
        }
    }
    if (C.midpnt > start)
        C.midpnt += tstep / C.m_speed;
    cil[c].midpnt = C.midpnt;
    cil[c].Tstroke = C.Tstroke;
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
    if ((e = upload(&s.x[0], pos3, nb * 3, st)) != hipSuccess || (e = upload(&s.x[1], pos3, nb * 3, st)) != hipSuccess ||
        (e = upload(&s.v, vel4, nb * 4, st)) != hipSuccess ||
        (e = upload(&s.f, (const double *)nullptr, nb * 3, st)) != hipSuccess ||
        (e = upload(&s.tstart, tstart.data(), tstart.size(), st)) != hipSuccess ||
        (e = upload(&s.terms, terms.data(), terms.size(), st)) != hipSuccess ||
        (e = upload(&s.cil, cil.data(), cil.size(), st)) != hipSuccess ||
        (e = upload(&s.newlen, newlen.data(), newlen.size(), st)) != hipSuccess ||
        (e = upload(&s.oldlen, oldlen.data(), oldlen.size(), st)) != hipSuccess)
        return e;
    if (mobile && (e = upload(&s.mobile, mobile, nb, st)) != hipSuccess)
        return e;
    if ((e = hipMemsetAsync(s.f, 0, sizeof(double) * 3 * (nb ? nb : 1), st)) != hipSuccess)
        return e;
    s.cur = 0;
    return hipSuccess;
}

void md_free(MdSystem &s)
{
    void *p[] = {s.x[0], s.x[1], s.v, s.f, s.mobile, s.tstart, s.terms, s.cil, s.newlen, s.oldlen};
    for (void *q : p)
        if (q)
            (void)hipFree(q);
    s = MdSystem{};
}

hipError_t md_run(MdSystem &s, double tstep, hipStream_t st)
{
    if (s.nb == 0)
        return hipSuccess;
    const int n = int((tstep + 0.00000001) / s.ststep);
    const double dt = s.ststep;
    const dim3 grid((s.nb + kMdBlock - 1) / kMdBlock), block(kMdBlock);
    const dim3 cgrid((s.ncil + 63) / 64 > 0 ? (s.ncil + 63) / 64 : 1), cblock(64);
    // the reference's globalt at each force evaluation m = 0 .. n: t0, t0,
    // then t0 + dt, t0 + 2 dt, ... by repeated addition (simulation.cpp:229-246)
    double g = s.globalt;
    auto recalc_kick = [&](double gt, double c, int drift) {
        if (s.ncil && std::fmod(gt + dt / 10., 0.01) < dt)
            hipLaunchKernelGGL(k_md_cilia_len, cgrid, cblock, 0, st, s.ncil, s.cil, s.x[s.cur], s.newlen, s.oldlen);
        hipLaunchKernelGGL(k_md_substep, grid, block, 0, st, s.nb, s.x[s.cur], s.x[1 - s.cur], s.v, s.f, s.mobile,
                           s.tstart, s.terms, s.cil, s.newlen, c, dt, drift);
        if (drift)
            s.cur = 1 - s.cur;
    };
    recalc_kick(g, dt / 2.0, 1);  // recalc, md_loop_start
    for (int k = 1; k < n; k++) {  // recalc at t0 (k = 1) or t0 + (k-1) dt, md_loop
        recalc_kick(g, dt, 1);
        g += dt;
    }
    recalc_kick(g, dt / 2.0, 0);  // last recalc, md_loop_stop (half kick)
    g += dt;
    if (s.ncil)
        hipLaunchKernelGGL(k_md_switch, cgrid, cblock, 0, st, s.ncil, s.cil, s.x[s.cur], g, tstep);
    s.globalt = g;
    return hipGetLastError();
}

}  // namespace mcm
