// Tile-fused MPCD step: the per-tile code of the FP32-bulk and FP64 forms.
//
// One workgroup of 256 threads per tile of TX x TY x TZ collision cells.
//   1. Every thread loads up to K particle records (home segment + the 26
//      inbound outboxes) and keeps the position part in registers; the
//      velocities go to LDS.
//   2. An LDS counting sort (count, exclusive scan, scatter of slot indices)
//      groups the tile's particles by collision cell.
//   3. One thread per cell walks its particles: sums, MBS thermostat,
//      rotation (mpcd.cpp:158-236), new velocities written back to LDS.
//   4. Every thread streams its particles with the new velocities, bins them
//      under the next shift, and writes each record to the new home segment
//      or to the outbox of the neighbouring tile it moved to.
// Every record is read from HBM once and written once per step.
//
// This header holds the tile code; the entry kernels and launchers live in
// mpcd_step_f32.hip, mpcd_step_f64.hip (the main pass over every tile, one
// box or an ensemble) and mpcd_step_later_f32.hip / mpcd_step_later_f64.hip
// (the bead pass and the large pass over the tiles the main pass left),
// compiled as separate translation units.
#pragma once

#include <type_traits>

#include "mpcd_kernels.hpp"

namespace mcm {

// ------------------------------------------------------------ helpers
template <int W>
__device__ __forceinline__ void tile_segments(const StepArgs &A, const TileCoord &tc, int *s_seg,
                                              const uint32_t **s_src, int *s_ocnt)
{
    const int tid = threadIdx.x;
    if (tid < 27) {
        int cnt;
        long long base;
        segment_info(A.g, tc, tid, A.in_home_cnt, A.in_ob_cnt, cnt, base);
        s_seg[tid + 1] = cnt;
        s_src[tid] = (tid == 0 ? A.in_home : A.in_ob) + base * W;
        s_ocnt[tid] = 0;
    }
}

// Thread 0: the prefix over the 27 segments and the tile's load.  A tile
// holding more than cap particles goes to the large pass (from the main and
// bead passes) or, in that pass, is flagged and truncated; in the main
// pass a tile holding beads (or one bead_tile names: cells a rigid sphere
// cuts) goes to the bead pass.  *s_defer tells every thread, after the
// next barrier, to leave the tile alone.
template <int PASS>
__device__ __forceinline__ void tile_admit(const StepArgs &A, int *s_seg, int cap, int tile, DiagSlot *D,
                                           int *s_defer, bool bead_tile = false)
{
    if (threadIdx.x == 0) {
        s_seg[0] = 0;
        for (int s = 1; s <= 27; ++s)
            s_seg[s] += s_seg[s - 1];
        int defer = 0;
        if (PASS == kPassMain)
            atomicMax(&D->max_load, s_seg[27]);
        if (s_seg[27] > cap) {
            const int k = PASS == kPassLarge ? kDeferCap : atomicAdd(&A.diag->defer_n, 1);
            if (k < kDeferCap) {
                A.diag->defer_tile[k] = tile;
                defer = 1;
            } else {
                atomicOr(A.flags, kFlagTileCap);
            }
        } else if (PASS == kPassMain && (bead_tile || (A.nb > 0 && A.b_tile_start[tile] < A.b_tile_start[tile + 1]))) {
            A.bead_tiles[atomicAdd(&A.diag->bead_n, 1)] = tile;
            defer = 1;
        }
        *s_defer = defer;
    }
}

// Beads of the tile: first sorted index and count per local cell.
__device__ __forceinline__ void tile_beads(const StepArgs &A, int t, int *s_bfirst, int *s_bn)
{
    const int b0 = A.b_tile_start[t], b1 = A.b_tile_start[t + 1];
    for (int i = b0 + (int)threadIdx.x; i < b1; i += kBlock) {
        const int lc = A.b_lc[i];
        if (i == b0 || A.b_lc[i - 1] != lc)
            s_bfirst[lc] = i;
        atomicAdd(&s_bn[lc], 1);
    }
}

// Loads the tile's records into rec[K][W]: the home segment (most of them)
// straight from its base, the rest after a search of the inbound outboxes
// that resumes, as k grows, where it stopped.
template <int K, int W>
__device__ __forceinline__ void load_tile(const int *s_seg, const uint32_t *const *s_src, int total,
                                          uint32_t (&rec)[K][W])
{
    const int tid = threadIdx.x;
    const int home_n = s_seg[1];
    const uint32_t *home = s_src[0];
    int s = 1;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int j = tid + k * kBlock;
        if (j < total) {
            const uint32_t *p;
            if (j < home_n) {
                p = home + __umul24((unsigned)j, (unsigned)W);
            } else {
                while (j >= s_seg[s + 1])
                    ++s;
                p = s_src[s] + __umul24((unsigned)(j - s_seg[s]), (unsigned)W);
            }
            ld_words<W>(p, rec[k]);
        }
    }
}

// Exclusive scan of NC <= 256 counts by one wave.
template <int NC>
__device__ __forceinline__ void wave0_exclusive_scan(const int *in, int *out)
{
    constexpr int PER = (NC + 63) / 64;
    const int l = lane_id();
    int v[PER];
    int sum = 0;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int idx = l * PER + i;
        v[i] = idx < NC ? in[idx] : 0;
        sum += v[i];
    }
    int incl = sum;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const int t = __shfl_up(incl, off);
        if (l >= off)
            incl += t;
    }
    int run = incl - sum;
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int idx = l * PER + i;
        if (idx < NC)
            out[idx] = run;
        run += v[i];
    }
}

// Where a tile's records go: segment 0 is its new home, 1 + d the outbox of
// direction d.  Filled once per workgroup, so the per-record path reads
// LDS rather than the kernel-argument tables.
struct DestTable {
    uint32_t *base[27];
    int cap[27];
};

template <int W>
__device__ __forceinline__ void tile_dests(const StepArgs &A, int t, DestTable &D)
{
    const int tid = threadIdx.x;
    if (tid < 27) {
        const Geom &g = A.g;
        if (tid == 0) {
            D.base[0] = A.out_home + (long long)t * g.cap_home * W;
            D.cap[0] = g.cap_home;
        } else {
            D.base[tid] = A.out_ob + ((long long)t * g.ob_stride + g.ob_off[tid - 1]) * W;
            D.cap[tid] = g.ob_cap[tid - 1];
        }
    }
}

// Writes one record to the tile's new home, an outbox, or the overflow list
// and returns where it went (null: no particle, or a full overflow list).
// dest: 0 home, 1 + d outbox d, 27 overflow, -1 no particle.  Every lane of
// the wave calls it (the home slot is a wave-aggregated bump).
template <int W>
__device__ __forceinline__ uint32_t *place_record(const StepArgs &A, const DestTable &D, int *s_ocnt, int dest,
                                                  const uint32_t (&r)[W])
{
    const int hr = wave_bump(&s_ocnt[0], dest == 0);
    if (dest < 0)
        return nullptr;
    uint32_t *dst = nullptr;
    if (dest < 27) {
        const int rk = dest == 0 ? hr : atomicAdd(&s_ocnt[dest], 1);
        if (rk < D.cap[dest])
            dst = D.base[dest] + __umul24((unsigned)rk, (unsigned)W);
    }
    if (!dst) {
        const int slot = atomicAdd(A.ovf_cnt, 1);
        if (slot < A.g.cap_ovf)
            dst = A.ovf + (long long)slot * W;
        else
            atomicOr(A.flags, kFlagOverflowCap);
    }
    if (dst)
        st_words<W>(dst, r);
    return dst;
}

__device__ __forceinline__ void write_counts(const StepArgs &A, int t, const int *s_ocnt)
{
    const int tid = threadIdx.x;
    if (tid < 27) {
        const int cnt = s_ocnt[tid];
        if (tid == 0)
            A.out_home_cnt[t] = min(cnt, A.g.cap_home);
        else
            A.out_ob_cnt[t * 26 + tid - 1] = min(cnt, A.g.ob_cap[tid - 1]);
    }
}

// Thermostat factor and rotation matrix of a cell from its Gamma variate G
// and axis (mpcd.cpp:204-212, 269-295).
__device__ __forceinline__ double scale_rotation(const StepArgs &A, double G, const double ax[3], int num, double rel,
                                                 DiagSlot *D, double R[9])
{
    double sc = 1.0;
    if (A.thermostat && rel > 0 && num > 1)
        sc = sqrt(A.kT * G / (rel * 0.5));
    else if (!(rel > 0 && num > 1) && !(rel * rel < 1e-8))
        atomicAdd((unsigned long long *)&D->rel_viol, 1ull);
    rotmat(ax, A.calpha, A.salpha, R);
    return sc;
}

// v' = u + sc R (v - u), in the reference's operation order (mpcd.cpp:214-231)
__device__ __forceinline__ void srd_rotate(const double R[9], double ux, double uy, double uz, double sc, double v0,
                                           double v1, double v2, double &n0, double &n1, double &n2)
{
    const double r0 = v0 - ux, r1 = v1 - uy, r2 = v2 - uz;
    n0 = ux + (R[0] * r0 + R[1] * r1 + R[2] * r2) * sc;
    n1 = uy + (R[3] * r0 + R[4] * r1 + R[5] * r2) * sc;
    n2 = uz + (R[6] * r0 + R[7] * r1 + R[8] * r2) * sc;
}

// ------------------------------------------------------------ rigid sphere
constexpr uint64_t kTagSphv = 0x53504856ull;

__device__ __forceinline__ double min_image(double d, int L) { return d - L * rint(d / L); }

// Candidate k of a cell's virtual sphere particles: offset u in the cell
// (words 0-2) and three normals (Box-Muller on words (3,4), both branches,
// and (5,6)), block {cell, step, k, 'SPHV'}.
__device__ __forceinline__ void sphv_draw(uint64_t seed, uint64_t step, uint64_t cell, uint64_t k, double u[3],
                                          double nr[3])
{
    const U64x4 o = philox4x64(cell, step, k, kTagSphv, seed, kKeyHi);
    u[0] = u32_open01((uint32_t)o.v[0]);
    u[1] = u32_open01((uint32_t)(o.v[0] >> 32));
    u[2] = u32_open01((uint32_t)o.v[1]);
    const double r1 = sqrt(-2.0 * log(u32_open01((uint32_t)(o.v[1] >> 32))));
    const double r2 = sqrt(-2.0 * log(u32_open01((uint32_t)(o.v[2] >> 32))));
    double sp, cp;
    sincospi(2.0 * u32_open01((uint32_t)o.v[2]), &sp, &cp);
    nr[0] = r1 * cp;
    nr[1] = r1 * sp;
    nr[2] = r2 * cospi(2.0 * u32_open01((uint32_t)o.v[3]));
}

// Velocity of the rigid-body point d (relative to the centre).
__device__ __forceinline__ void sphere_point_velocity(const SphereDev *S, const double d[3], double v[3])
{
    v[0] = S->V[0] + (S->W[1] * d[2] - S->W[2] * d[1]);
    v[1] = S->V[1] + (S->W[2] * d[0] - S->W[0] * d[2]);
    v[2] = S->V[2] + (S->W[0] * d[1] - S->W[1] * d[0]);
}

// Virtual particle q of a cell cut by the sphere (dcen: the cell centre
// relative to the sphere's, minimum image): in = false when the candidate
// point lies outside; else its offset d from the centre and its velocity
// (rigid-body motion plus sqrt(kT) noise).
struct SphereVirtual {
    double d0, d1, d2, v0, v1, v2;
    bool in;
};

__device__ __forceinline__ SphereVirtual sphere_virtual(const StepArgs &A, long long gid, double c0, double c1,
                                                        double c2, int q)
{
    const SphereDev *S = A.sphere;
    double u[3], nr[3];
    sphv_draw(A.seed, A.step, (uint64_t)gid, (uint64_t)q, u, nr);
    if (A.inj_sph) {
        const long long at = 3 * ((long long)A.rho * gid + q);
        nr[0] = A.inj_sph[at];
        nr[1] = A.inj_sph[at + 1];
        nr[2] = A.inj_sph[at + 2];
    }
    SphereVirtual r;
    const double d[3] = {c0 + (u[0] - 0.5), c1 + (u[1] - 0.5), c2 + (u[2] - 0.5)};
    r.in = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] < S->R * S->R;
    double v[3];
    sphere_point_velocity(S, d, v);
    const double sk = sqrt(A.kT);
    r.d0 = d[0];
    r.d1 = d[1];
    r.d2 = d[2];
    r.v0 = v[0] + sk * nr[0];
    r.v1 = v[1] + sk * nr[1];
    r.v2 = v[2] + sk * nr[2];
    return r;
}

// Bounce-back of one particle streamed from xo to x = xo + n h off the
// sphere, which moves from X to X + V h during the step: first crossing of
// |d| = R in the relative motion, then no-slip (n' = 2 u_s - n) or slip
// (the normal relative velocity flips), and the remaining time with n'.
// The impulse n - n' and its moment go to J, T.
__device__ __forceinline__ void sphere_bounce(const StepArgs &A, const double xo[3], double x[3], double n[3],
                                              double J[3], double T[3])
{
    const SphereDev *S = A.sphere;
    const double ts = A.h, R2 = S->R * S->R;
    double d1[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
        d1[k] = min_image(x[k] - (S->X[k] + S->V[k] * ts), A.g.L[k]);
    if (d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2] >= R2)
        return;
    double w[3], d0[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        w[k] = n[k] - S->V[k];
        d0[k] = d1[k] - w[k] * ts;
    }
    const double qa = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
    const double qb = d0[0] * w[0] + d0[1] * w[1] + d0[2] * w[2];
    const double qc = d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2] - R2;
    double t = 0.0;
    if (qc > 0.0 && qa > 0.0) {
        const double disc = qb * qb - qa * qc;
        t = (-qb - sqrt(disc > 0.0 ? disc : 0.0)) / qa;
        t = t < 0.0 ? 0.0 : (t > ts ? ts : t);
    }
    double dc[3], us[3], vn[3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
        dc[k] = d0[k] + w[k] * t;
    sphere_point_velocity(S, dc, us);
    if (S->slip) {
        const double r = sqrt(dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]);
        const double nh[3] = {dc[0] / r, dc[1] / r, dc[2] / r};
        const double vr = (n[0] - us[0]) * nh[0] + (n[1] - us[1]) * nh[1] + (n[2] - us[2]) * nh[2];
#pragma unroll
        for (int k = 0; k < 3; ++k)
            vn[k] = n[k] - 2.0 * vr * nh[k];
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k)
            vn[k] = 2.0 * us[k] - n[k];
    }
    double Jp[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        x[k] = (xo[k] + n[k] * t) + vn[k] * (ts - t);
        Jp[k] = n[k] - vn[k];
        J[k] += Jp[k];
        n[k] = vn[k];
    }
    T[0] += dc[1] * Jp[2] - dc[2] * Jp[1];
    T[1] += dc[2] * Jp[0] - dc[0] * Jp[2];
    T[2] += dc[0] * Jp[1] - dc[1] * Jp[0];
}

// Whether a particle streamed to x ends inside the sphere (at X + V h).
__device__ __forceinline__ bool sphere_inside(const StepArgs &A, double x, double y, double z)
{
    const SphereDev *S = A.sphere;
    const double ts = A.h;
    const double d0 = min_image(x - (S->X[0] + S->V[0] * ts), A.g.L[0]);
    const double d1 = min_image(y - (S->X[1] + S->V[1] * ts), A.g.L[1]);
    const double d2 = min_image(z - (S->X[2] + S->V[2] * ts), A.g.L[2]);
    return d0 * d0 + d1 * d1 + d2 * d2 < S->R * S->R;
}

// The bounce of one particle that ended inside the sphere.  Impulse,
// moment and the change of the diagnostics go to the tile's LDS
// accumulators.
struct SphereHit {
    double x, y, z, n0, n1, n2;
};

__device__ __forceinline__ SphereHit sphere_hit(const StepArgs &A, double xo0, double xo1, double xo2, double x, double y,
                                             double z, double n0, double n1, double n2, double *s_sph, double *s_dred)
{
    const double xo[3] = {xo0, xo1, xo2};
    double xs[3] = {x, y, z}, n[3] = {n0, n1, n2};
    double J[3] = {0.0, 0.0, 0.0}, T[3] = {0.0, 0.0, 0.0};
    sphere_bounce(A, xo, xs, n, J, T);
    for (int q = 0; q < 3; ++q) {
        atomicAdd(&s_sph[q], J[q]);
        atomicAdd(&s_sph[3 + q], T[q]);
    }
    atomicAdd(&s_dred[1], (n[0] * n[0] + n[1] * n[1] + n[2] * n[2]) - (n0 * n0 + n1 * n1 + n2 * n2));
    atomicAdd(&s_dred[2], n[0] - n0);
    atomicAdd(&s_dred[3], n[1] - n1);
    atomicAdd(&s_dred[4], n[2] - n2);
    return SphereHit{xs[0], xs[1], xs[2], n[0], n[1], n[2]};
}

// --------------------------------------- rigid sphere, FP32-bulk records
// The FP32-bulk form takes the sphere as the FP64 form does: the same
// virtual particles (sphere_virtual) in every cell it cuts, the same
// bounce-back (sphere_bounce) in the stream.  A cell it cuts has no fluid
// owner of the residual: the sphere takes the cell's residual sum instead
// (so fluid + sphere momentum is conserved to FP64 roundoff, the fluid's
// in exact integers), and the virtual particles' angular impulse.  Tiles
// holding such cells go to the bead pass (FP64 map there).

// Whether the sphere cuts collision cell (cx, cy, cz) of the current
// frame; dcen receives the cell centre relative to the sphere's centre
// (minimum image), as in the FP64 form.
__device__ __forceinline__ bool sphere_cuts_cell(const StepArgs &A, int cx, int cy, int cz, double (&dcen)[3])
{
    const int cc[3] = {cx, cy, cz};
    double q2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        dcen[k] = min_image((cc[k] + 0.5 - A.a[k]) - A.sphere->X[k], A.g.L[k]);
        const double q = fabs(dcen[k]) - 0.5;
        if (q > 0.0)
            q2 += q * q;
    }
    return q2 < A.sphere->R * A.sphere->R;
}

// Squared distance from point P to the absolute region of the tile's cells
// (current frame) grown by m on every side, minimum image per axis.
__device__ __forceinline__ double tile_dist2(const StepArgs &A, const TileCoord &tc, const int T[3], const double P[3],
                                             double m)
{
    const int o[3] = {tc.ox, tc.oy, tc.oz};
    double q2 = 0.0;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double c = (o[k] + 0.5 * T[k]) - A.a[k];
        const double q = fabs(min_image(P[k] - c, A.g.L[k])) - (0.5 * T[k] + m);
        if (q > 0.0)
            q2 += q * q;
    }
    return q2;
}

// A particle that moves at most kSolidReach cells in a step can reach a
// solid (the sphere, a wall) only from a tile that close to it; a faster
// one is tested wherever it is.
constexpr double kSolidReach = 2.0;

// Destination of a record whose new cell is nc (frame a_next): 0 home,
// 1 + d outbox d, 27 overflow list.
template <int TX, int TY, int TZ>
__device__ __forceinline__ int dest_of_cell(const Geom &g, const TileCoord &tc, const int (&nc)[3])
{
    const int tt[3] = {tc.tx, tc.ty, tc.tz};
    constexpr int T[3] = {TX, TY, TZ};
    int st[3];
    bool far = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        st[a] = tile_dir(tt[a], nc[a] / T[a], g.nt[a]);
        far |= st[a] == 2;
    }
    return far ? 27 : ((st[0] | st[1] | st[2]) == 0 ? 0 : 1 + dir26(st[0], st[1], st[2]));
}

// A bounced FP32-bulk particle at absolute position x with velocity n
// (FP64) encoded again as the form encodes any state: the position in
// the cell and offsets of frame a_next, the velocity on the 2^-21 grid
// (wn).  The exact changes of the fluid's momentum (2^-21 units) and |q|^2
// (2^-42 units) against the velocity w it streamed with go to dm and de.
// Returns the destination of the record r.
template <int TX, int TY, int TZ, int W>
__device__ __forceinline__ int reencode_f32(const StepArgs &A, const TileCoord &tc, const double (&x)[3],
                                            const double (&n)[3], const float (&w)[3], float (&wn)[3],
                                            uint32_t (&r)[W], long long (&dm)[3], long long &de, bool &bad_v)
{
    int nc[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        wn[k] = quant_v(n[k]);
        bad_v |= !(fabsf(wn[k]) < 512.0f);
        const long long q0 = q32_of(w[k]), q1 = q32_of(wn[k]);
        dm[k] += q1 - q0;
        de += q1 * q1 - q0 * q0;
        const double sh = x[k] + A.a_next[k];
        const double f = floor(sh);
        int c = box_axis(sh, A.g.L[k], A.g.invL[k]);
        float o = (float)(sh - f);
        if (o >= 1.0f) {
            o = 0.0f;
            c = c + 1 == A.g.L[k] ? 0 : c + 1;
        }
        nc[k] = c;
        r[k] = fbits(o);
        r[3 + k] = fbits(wn[k]);
    }
    r[6] = pack_cell(nc[0], nc[1], nc[2]);
    return dest_of_cell<TX, TY, TZ>(A.g, tc, nc);
}

// The bounce-back of an FP32-bulk particle streamed with w from record rec
// (current frame) to the new record r (frame a_next): when it ends inside
// the sphere, sphere_bounce reflects it in FP64 and r is encoded again; the
// impulse, w minus the reflected grid velocity (exact), and its moment go
// to s_sph.  Returns the destination of the record.
template <int TX, int TY, int TZ, int W>
__device__ __noinline__ int sphere_bounce_f32(const StepArgs &A, const TileCoord &tc, const uint32_t (&rec)[W],
                                                 const float (&w)[3], uint32_t (&r)[W], int dest, double *s_sph,
                                                 long long (&dm)[3], long long &de, bool &bad_v)
{
    const uint32_t key = rec[6];
    const int cc[3] = {(int)(key & 1023u), (int)((key >> 10) & 1023u), (int)(key >> 20)};
    double xo[3], x[3], n[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        xo[k] = ((double)cc[k] - A.a[k]) + (double)bitsf(rec[k]);
        n[k] = w[k];
        x[k] = xo[k] + n[k] * A.h;
    }
    if (!sphere_inside(A, x[0], x[1], x[2]))
        return dest;
    double J[3] = {0.0, 0.0, 0.0}, T[3] = {0.0, 0.0, 0.0};
    sphere_bounce(A, xo, x, n, J, T);
    float wn[3];
    dest = reencode_f32<TX, TY, TZ, W>(A, tc, x, n, w, wn, r, dm, de, bad_v);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        atomicAdd(&s_sph[k], (double)w[k] - (double)wn[k]);
        atomicAdd(&s_sph[3 + k], T[k]);
    }
    return dest;
}

// The planar walls of the FP32-bulk form (z = 0 and z = L_z - 1, the
// reference's SLIP / NOSLIP build): a particle streamed with w from record
// rec whose height leaves [0, L_z - 1] bounces as in MPCD::move
// (wall_bounce, FP64, its height taken in [0, L_z) before the step) and r
// is encoded again.  Returns the destination of the record.
template <int TX, int TY, int TZ, int W>
__device__ __noinline__ int wall_bounce_f32(const StepArgs &A, const TileCoord &tc, const uint32_t (&rec)[W],
                                               const float (&w)[3], uint32_t (&r)[W], int dest, long long (&dm)[3],
                                               long long &de, bool &bad_v)
{
    const uint32_t key = rec[6];
    const int cc[3] = {(int)(key & 1023u), (int)((key >> 10) & 1023u), (int)(key >> 20)};
    const double Lz = A.g.L[2], ztop = Lz - 1.0;
    double x[3], n[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        double xo = ((double)cc[k] - A.a[k]) + (double)bitsf(rec[k]);
        if (k == 2)
            xo -= Lz * floor(xo / Lz);
        n[k] = w[k];
        x[k] = xo + n[k] * A.h;
    }
    if (!(x[2] > ztop || x[2] < 0.0))
        return dest;
    wall_bounce(A.walls, ztop, A.vw[0], A.vw[1], x[0], x[1], x[2], n[0], n[1], n[2]);
    float wn[3];
    return reencode_f32<TX, TY, TZ, W>(A, tc, x, n, w, wn, r, dm, de, bad_v);
}

// The bounce of a streamed particle off the tile code's solid object OBJ
// (the bounces themselves out of line: they are rare, and the step's
// unrolled particle loop keeps its registers).
template <int OBJ, int TX, int TY, int TZ, int W>
__device__ __forceinline__ int solid_bounce_f32(const StepArgs &A, const TileCoord &tc, const uint32_t (&rec)[W],
                                                const float (&w)[3], uint32_t (&r)[W], int dest, double *s_sph,
                                                long long (&dm)[3], long long &de, bool &bad_v)
{
    if constexpr (OBJ == kObjSphere)
        return sphere_bounce_f32<TX, TY, TZ, W>(A, tc, rec, w, r, dest, s_sph, dm, de, bad_v);
    else
        return wall_bounce_f32<TX, TY, TZ, W>(A, tc, rec, w, r, dest, dm, de, bad_v);
}

// Whether a particle streamed with w may reach a solid from a tile that is
// not near one (it moves more than kSolidReach cells, or w is not finite).
__device__ __forceinline__ bool solid_fast(const StepArgs &A, const float (&w)[3])
{
    return !(fmaxf(fmaxf(fabsf(w[0]), fabsf(w[1])), fabsf(w[2])) * A.hf <= (float)kSolidReach);
}

// ===================================================================== FP32
// Record (7 words, 8 with ids): offsets ox oy oz in [0,1) of the shifted
// cell, velocities on the 2^-21 grid, packed cell key, [particle id].
//
// Every per-cell reduction is an exact integer sum (q = v 2^21, |q|^2 split
// at bit 40) or a maximum, so its result does not depend on the order the
// tile's particles are held in.  Each thread keeps its K records in
// registers; only the velocities visit LDS, in cell order, where one thread
// per cell sums them serially (no LDS atomics on the sums).
//   pass 1  per particle: rank in the cell (the only LDS atomic), then the
//           velocity goes to its cell-ordered slot
//   pass 2  per cell: sums, owner key, u, relative energy, thermostat,
//           rotation -> the affine map v' = M v + b (M = sc R, b = u - M u)
//   pass 3  per particle: v' on the grid into the slot; every particle but
//           the cell's residual owner streams and is written at once
//   pass 4  per cell: the residual sum (q - q'); the owner of each fluid-only
//           cell adds it (momentum exact) and is written
// The owner is the particle with the largest content key (velocity bits
// before the collision; ties, which need equal bits, go on to the record
// after it and the id), so the choice is the same in any order.

// Local cell of a record's packed key; flags keys outside the tile.
template <int TX, int TY, int TZ>
__device__ __forceinline__ int local_cell(const StepArgs &A, const TileCoord &tc, uint32_t key)
{
    const int lx = (int)(key & 1023u) - tc.ox;
    const int ly = (int)((key >> 10) & 1023u) - tc.oy;
    const int lz = (int)(key >> 20) - tc.oz;
    if ((unsigned)lx >= (unsigned)TX || (unsigned)ly >= (unsigned)TY || (unsigned)lz >= (unsigned)TZ) {
        atomicOr(A.flags, kFlagMisbinned);
        return 0;
    }
    return lx + TX * (ly + TY * lz);
}

__device__ __forceinline__ unsigned long long key2w(uint32_t hi, uint32_t lo)
{
    return ((unsigned long long)hi << 32) | lo;
}

// Streams one FP32-bulk record with velocity w (off + w h + (a_next - a) per
// axis in FP32), bins it under the next shift and returns its destination
// (0 home, 1 + d outbox d, 27 overflow list); r receives the new record.
template <int TX, int TY, int TZ, int W>
__device__ __forceinline__ int stream_rebin(const StepArgs &A, const TileCoord &tc, const uint32_t (&rec)[W],
                                            const float (&w)[3], uint32_t (&r)[W], bool &bad_v)
{
    const Geom &g = A.g;
    const uint32_t key = rec[6];
    const int cc[3] = {(int)(key & 1023u), (int)((key >> 10) & 1023u), (int)(key >> 20)};
    const int tt[3] = {tc.tx, tc.ty, tc.tz};
    int nc[3], st[3];
    bool far = false;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        bad_v |= !(fabsf(w[a]) < 512.0f);
        const float t1 = __fmul_rn(w[a], A.hf);
        const float t2 = __fadd_rn(bitsf(rec[a]), t1);
        const float pp = __fadd_rn(t2, A.da[a]);
        const float fl = floorf(pp);
        float o = __fsub_rn(pp, fl);
        const bool up = o >= 1.0f;
        o = up ? 0.0f : o;
        r[a] = fbits(o);
        r[3 + a] = fbits(w[a]);
        const int L = g.L[a];
        int cn = cc[a] + (int)fl + (int)up;
        cn += cn < 0 ? L : 0;
        cn -= cn >= L ? L : 0;
        if ((unsigned)cn >= (unsigned)L) {  // moved more than a box length: flagged (the step fails)
            bad_v = true;
            cn = min(max(cn, 0), L - 1);     // any cell of the box: the record stays in bounds
        }
        nc[a] = cn;
        const unsigned Ta = a == 0 ? TX : (a == 1 ? TY : TZ);
        st[a] = tile_dir(tt[a], (int)((unsigned)cn / Ta), g.nt[a]);
        far |= st[a] == 2;
    }
    r[6] = pack_cell(nc[0], nc[1], nc[2]);
    if constexpr (W == 8)
        r[7] = rec[7];
    return far ? 27 : ((st[0] | st[1] | st[2]) == 0 ? 0 : 1 + dir26(st[0], st[1], st[2]));
}

// Whether every move of at most one tile from this tile stays inside the
// box without a periodic wrap, along every axis, and the two neighbours of
// every axis are distinct, full tiles: then stream_rebin_interior applies.
__device__ __forceinline__ bool tile_interior(const Geom &g, const TileCoord &tc)
{
    const int tt[3] = {tc.tx, tc.ty, tc.tz};
    bool in = true;
#pragma unroll
    for (int a = 0; a < 3; ++a)
        in = in && tt[a] >= 1 && (tt[a] + 2) * g.T[a] <= g.L[a];
    return in;
}

// stream_rebin for a particle of an interior tile (tile_interior): the same
// stream arithmetic, then the move in cells read against the tile's local
// coordinates lc (no wrap is possible); the packed cell key moves by the
// same cells.  With power-of-two tile sides, (local + move + T) / T is the
// tile step plus one (0, 1, 2) of an axis, and the three of them index the
// 27 directions directly (13: home).  A move of more than one tile goes to
// stream_rebin.
template <int TX, int TY, int TZ, int W>
__device__ __forceinline__ int stream_rebin_interior(const StepArgs &A, const TileCoord &tc, int lc,
                                                     const uint32_t (&rec)[W], const float (&w)[3],
                                                     uint32_t (&r)[W], bool &bad_v)
{
    static_assert((TX & (TX - 1)) == 0 && (TY & (TY - 1)) == 0 && (TZ & (TZ - 1)) == 0, "power-of-two tile sides");
    constexpr int T[3] = {TX, TY, TZ};
    constexpr int SH[3] = {__builtin_ctz(TX), __builtin_ctz(TY), __builtin_ctz(TZ)};
    const int l[3] = {lc & (TX - 1), (lc >> SH[0]) & (TY - 1), lc >> (SH[0] + SH[1])};
    // |w| < 512 on all three axes, NaN included: the magnitude bits as integers
    const uint32_t wmax = max(max(fbits(w[0]) & 0x7fffffffu, fbits(w[1]) & 0x7fffffffu), fbits(w[2]) & 0x7fffffffu);
    bad_v |= wmax >= 0x44000000u;  // 512.0f
    int di[3];
    uint32_t sx[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float t1 = __fmul_rn(w[a], A.hf);
        const float t2 = __fadd_rn(bitsf(rec[a]), t1);
        const float pp = __fadd_rn(t2, A.da[a]);
        const float fl = floorf(pp);
        float o = __fsub_rn(pp, fl);
        const bool up = o >= 1.0f;
        o = up ? 0.0f : o;
        r[a] = fbits(o);
        r[3 + a] = fbits(w[a]);
        di[a] = (int)fl + (int)up;
        sx[a] = (uint32_t)(l[a] + di[a] + T[a]) >> SH[a];  // 0, 1, 2; anything else: farther
    }
    if (max(max(sx[0], sx[1]), sx[2]) > 2u)
        return stream_rebin<TX, TY, TZ, W>(A, tc, rec, w, r, bad_v);
    r[6] = rec[6] + (uint32_t)(di[0] + (di[1] << 10) + (di[2] << 20));
    if constexpr (W == 8)
        r[7] = rec[7];
    const int i = (int)(sx[0] * 9u + sx[1] * 3u + sx[2]);
    return i == 13 ? 0 : (i < 13 ? i + 1 : i);  // home, else 1 + dir26
}

// Owner tie-break key of stage s >= 1 from a record whose velocity words
// already hold the new velocity.
template <int W>
__device__ __forceinline__ unsigned long long owner_stage_key(int s, const uint32_t (&rec)[W])
{
    if (s == 1)
        return key2w(rec[3], rec[4]);
    if (s == 2)
        return key2w(rec[5], rec[0]);
    if (s == 3)
        return key2w(rec[1], rec[2]);
    if constexpr (W == 8)
        return (unsigned long long)rec[7];
    return 0;
}

// Sum over the wave (every lane active), by the device library's DPP
// reduction: no LDS round trips.
extern "C" __device__ __attribute__((const)) unsigned long long __ockl_wfred_add_u64(unsigned long long);
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) { return __ockl_wfred_add_u64(v); }

// The velocity the residual owner keeps: its new grid velocity w plus the
// cell's residual sum dq (units of 2^-21).  Beyond the grid's exact range
// the rounding shows in the momentum account acc_m.
__device__ __forceinline__ float owner_velocity(float w, int dq, long long &acc_m)
{
    const int qf = (int)((unsigned)q32_of(w) + (unsigned)dq);
    const float wf = (float)((double)qf * kQinv);
    const int qa = q32_of(wf);
    if (qa != qf)
        acc_m += qa - qf;
    return wf;
}

// The owner's velocity: its new grid velocity w plus the cell's residual
// sum dq per axis; the change of its |q|^2 goes to de, any rounding beyond
// the grid's exact range to acc_m.
__device__ __forceinline__ void owner_take_residual(const float (&w)[3], const int *dq, float (&wf)[3],
                                                    long long &de, long long (&acc_m)[3])
{
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        wf[i] = owner_velocity(w[i], dq[i], acc_m[i]);
        const long long q0 = q32_of(w[i]), qa = q32_of(wf[i]);
        de += qa * qa - q0 * q0;
    }
}

// A deferred record of a tied cell (velocity words already new): it
// streams with that velocity; when it is the owner it keeps the velocity
// plus the cell's residual sum, whose change of |q|^2 goes to acc_e1 (the
// cell's sum after the collision already holds the record's new velocity).
// Returns its destination; r receives the new record.
template <int TX, int TY, int TZ, int W, bool TRACK>
__device__ __forceinline__ int finish_deferred(const StepArgs &A, const TileCoord &tc, const uint32_t (&q)[W],
                                               const int *dq, bool is_owner, unsigned long long &acc_e1,
                                               long long (&acc_m)[3], uint32_t (&r)[W], bool &bad_v)
{
    const float w[3] = {bitsf(q[3]), bitsf(q[4]), bitsf(q[5])};
    // a record that is not the owner keeps w: owner_velocity(w, 0) == w
    float wf[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        wf[i] = owner_velocity(w[i], is_owner ? dq[i] : 0, acc_m[i]);
        const long long q0 = q32_of(w[i]), qa = q32_of(wf[i]);
        acc_e1 += (unsigned long long)(qa * qa - q0 * q0);
    }
    if constexpr (TRACK) {
        if (A.dbg_cell) {
            const uint32_t key = q[6];
            A.dbg_cell[q[7]] =
                (int)(key & 1023u) + A.g.L[0] * ((int)((key >> 10) & 1023u) + A.g.L[1] * (int)(key >> 20));
            A.dbg_rank[q[7]] = -1;
        }
    }
    const int dest = stream_rebin<TX, TY, TZ, W>(A, tc, q, w, r, bad_v);
#pragma unroll
    for (int i = 0; i < 3; ++i)
        r[3 + i] = fbits(wf[i]);
    return dest;
}

// finish_deferred with a solid object OBJ: the record streams with its new
// velocity and bounces off the object if it reaches it; the owner then
// keeps the (reflected) velocity plus the residual.
template <int OBJ, int TX, int TY, int TZ, int W, bool TRACK>
__device__ __forceinline__ int finish_deferred_obj(const StepArgs &A, const TileCoord &tc, const uint32_t (&q)[W],
                                                   const int *dq, bool is_owner, unsigned long long &acc_e1,
                                                   long long (&acc_m)[3], uint32_t (&r)[W], bool &bad_v,
                                                   bool near, double *s_sph, long long (&sph_dm)[3],
                                                   long long &sph_de)
{
    const float w[3] = {bitsf(q[3]), bitsf(q[4]), bitsf(q[5])};
    if constexpr (TRACK) {
        if (A.dbg_cell) {
            const uint32_t key = q[6];
            A.dbg_cell[q[7]] =
                (int)(key & 1023u) + A.g.L[0] * ((int)((key >> 10) & 1023u) + A.g.L[1] * (int)(key >> 20));
            A.dbg_rank[q[7]] = -1;
        }
    }
    int dest = stream_rebin<TX, TY, TZ, W>(A, tc, q, w, r, bad_v);
    if (near || solid_fast(A, w))
        dest = solid_bounce_f32<OBJ, TX, TY, TZ, W>(A, tc, q, w, r, dest, s_sph, sph_dm, sph_de, bad_v);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const float wb = bitsf(r[3 + i]);
        const float wf = owner_velocity(wb, is_owner ? dq[i] : 0, acc_m[i]);
        const long long q0 = q32_of(wb), qa = q32_of(wf);
        acc_e1 += (unsigned long long)(qa * qa - q0 * q0);
        r[3 + i] = fbits(wf);
    }
    return dest;
}

// ------------------------------------------- MPC-AT in the FP32-bulk form
// The fluid's random velocities lie on the 2^-21 velocity grid and its
// lever arms on the 2^-24 grid of the cell offsets, so every fluid sum of
// the cell (random velocities, offsets, r x (v - xi), r r^T) is an exact
// integer and the collision does not depend on the order of the cell's
// particles.  A fluid particle's normals come from Philox4x32-10 of counter
// {cell, bits(ox), bits(oy), bits(oz)} (its offsets name it in its cell), a
// bead's from {cell, 0xffffffff, bead, 0xffffffff}, key {low word of the
// seed ^ 'F3AT', high word of the seed ^ low word of the step}: a
// Box-Muller pair from words 0 and 1 and the cosine branch of words 2 and
// 3, in FP32 on the transcendental unit (24-bit uniforms, as the cell
// draws).
constexpr uint64_t kTagF32At = 0x46334154ull;
constexpr double kRq = 16777216.0;                  // lever-arm grid 2^-24
constexpr double kRqinv = 1.0 / 16777216.0;
constexpr double kRQinv = 1.0 / 35184372088832.0;   // 2^-45: offset x velocity units
constexpr double kRRinv = 1.0 / 281474976710656.0;  // 2^-48: offset x offset units

__device__ __forceinline__ double local_coord(double x, double a);
__device__ __forceinline__ void at_omega(const double I[6], const double L[3], double w[3]);

__device__ __forceinline__ void at_f32_normals(uint64_t seed, uint64_t step, uint64_t cell, uint32_t c1, uint32_t c2,
                                               uint32_t c3, double n[3])
{
    const U32x4 w = philox4x32((uint32_t)cell, c1, c2, c3, (uint32_t)seed ^ (uint32_t)kTagF32At,
                               (uint32_t)(seed >> 32) ^ (uint32_t)step);
    constexpr float kLn2 = 0.69314718055994531f;
    const float r0 = __builtin_amdgcn_sqrtf(-2.0f * (__builtin_amdgcn_logf(u24_open01(w.v[0])) * kLn2));
    const float p0 = u24_open01(w.v[1]);
    const float r1 = __builtin_amdgcn_sqrtf(-2.0f * (__builtin_amdgcn_logf(u24_open01(w.v[2])) * kLn2));
    n[0] = (double)(r0 * __builtin_amdgcn_cosf(p0));
    n[1] = (double)(r0 * __builtin_amdgcn_sinf(p0));
    n[2] = (double)(r1 * __builtin_amdgcn_cosf(u24_open01(w.v[3])));
}

// A fluid record's random velocity in grid units (sqrt(kT) N(0,1)^3 2^21,
// rounded); injected normals are indexed by particle id.
template <int W, bool TRACK>
__device__ __forceinline__ void at_f32_fluid_xi(const StepArgs &A, uint64_t gid, const uint32_t (&rec)[W], int xq[3])
{
    double n[3];
    bool inj = false;
    if constexpr (TRACK) {
        if (A.inj_ran) {
            const long long id = (long long)rec[W - 1];
            n[0] = A.inj_ran[3 * id];
            n[1] = A.inj_ran[3 * id + 1];
            n[2] = A.inj_ran[3 * id + 2];
            inj = true;
        }
    }
    if (!inj)
        at_f32_normals(A.seed, A.step, gid, rec[0], rec[1], rec[2], n);
    const double sk = sqrt(A.kT);
#pragma unroll
    for (int a = 0; a < 3; ++a)
        xq[a] = (int)rint(sk * n[a] * kQ);
}

// Bead i (binned order) of the cell: random velocity sqrt(kT/m) N(0,1)^3 in
// FP64; injected normals at 3 (fluid count + bead).
__device__ __forceinline__ void at_f32_bead_xi(const StepArgs &A, uint64_t gid, int i, double xi[3])
{
    const int bi = A.b_idx[i];
    const double m = A.b_vel4[4 * (long long)i + 3];
    double n[3];
    if (A.inj_ran) {
        const long long at = 3 * (A.nf_total + (long long)bi);
        n[0] = A.inj_ran[at];
        n[1] = A.inj_ran[at + 1];
        n[2] = A.inj_ran[at + 2];
    } else {
        at_f32_normals(A.seed, A.step, gid, 0xffffffffu, (uint32_t)bi, 0xffffffffu, n);
    }
    const double sk = sqrt(A.kT / m);
#pragma unroll
    for (int a = 0; a < 3; ++a)
        xi[a] = sk * n[a];
}

// Lever arm of a record on the 2^-24 grid (an integer), from its offset.
__device__ __forceinline__ long long at_f32_rq(uint32_t off_bits) { return (long long)rint((double)bitsf(off_bits) * kRq); }

// A cell's draws of the FP32-bulk form before its particle count is known:
// the rotation axis, the first two Gamma candidates' normals and their
// uniforms as 32-bit words.
struct CellDrawsF32 {
    double ax[3];
    float x0, x1;
    uint32_t uw0, uw1;
};

template <int TX, int TY, int TZ, int K, bool TRACK, bool BEADS, int PASS, int RULE = kRuleSrd, int OBJ = kObjNone>
__device__ __forceinline__ void step_f32_tile(const StepArgs &A, int tile)
{
    constexpr bool SPHERE = OBJ == kObjSphere, WALLS = OBJ == kObjWalls, SOLID = OBJ != kObjNone;
    static_assert(!SOLID || RULE == kRuleSrd, "the solid objects take the SRD rule");
    constexpr int NC = TX * TY * TZ;
    static_assert(NC <= kBlock && NC <= 256, "one thread per cell, 8-bit local cells");
    static_assert(K <= 32, "deferral masks are 32-bit");
    constexpr int W = TRACK ? 8 : 7;
    constexpr int CAP = K * kBlock;
    static_assert(CAP <= 65536, "16-bit slots");
    constexpr int NB = BEADS ? NC : 1;
    constexpr int NSTAGE = TRACK ? 5 : 4;  // owner key stages after the first
    constexpr bool AT = RULE != kRuleSrd;     // MPC-AT (+a): the map is v' = b + xi + omega x r
    constexpr bool ATA = RULE == kRuleAtA;
    const Geom &g = A.g;
    const int tid = threadIdx.x;
    const TileCoord tc = tile_coord(g, tile);
    DiagSlot *D = &A.diag->slot[tc.t & (kDiagSlots - 1)];

    __shared__ int s_seg[28];
    __shared__ const uint32_t *s_src[27];
    __shared__ int s_ocnt[27];
    __shared__ DestTable s_dst;
    __shared__ int s_cnt[NC];
    __shared__ int s_start[NC];
    __shared__ float s_w[CAP][3];                // velocities in cell order: old, then new
    __shared__ int s_owner[NC];                  // slot of the cell's residual owner, -1: tied or none
    __shared__ uint32_t s_orec[NC][W + 1];       // the owner's streamed record and its destination
    __shared__ unsigned long long s_k1[NC];      // owner key, stage 0 (tied cells)
    __shared__ int s_claim[NC];
    __shared__ int s_dq[NC][3];                  // residual sum (q - q') of the cell
    // M (row-major), b of each cell in FP32; after pass 3 the same words hold
    // the maxima of the owner key stages of tied cells
    __shared__ __align__(16) float s_map[NC][12];
    static_assert(sizeof(float) * 12 * NC >= sizeof(unsigned long long) * (NSTAGE - 1) * NC, "stage keys fit");
    unsigned long long(*s_kst)[NC] = reinterpret_cast<unsigned long long(*)[NC]>(&s_map[0][0]);
    __shared__ int s_bfirst[NB];
    __shared__ int s_bn[NB];
    __shared__ double s_map64[NB][12];           // bead cells: the map in FP64
    __shared__ unsigned long long s_red[6];      // |q|^2 before, after; momentum xyz; |q|^2 change by the sphere
    __shared__ double s_dred[1];                 // bead energy before - after
    // solid objects: the sphere's impulse and moment from this tile; whether
    // streamed particles may reach the object; which cells hold its virtual
    // particles (bead pass)
    constexpr int NS = (SOLID && BEADS) ? NC : 1;
    __shared__ double s_sph[6];
    __shared__ int s_near;
    __shared__ int s_cut[NS];
    __shared__ int s_anytie;
    // candidates of tied cells held back in LDS (more go to the register path)
    constexpr int QCAP = 32;
    __shared__ uint32_t s_q[QCAP][W + 1];        // record (new velocity), local cell
    // the cell draws (axis, Gamma candidates) live in the cell's map row
    // until the cell's own thread replaces them with the map
    static_assert(sizeof(float) * 12 >= sizeof(CellDrawsF32), "cell draws fit a map row");
    __shared__ int s_qn;
    __shared__ int s_defer;
    // MPC-AT: each slot's random velocity (grid units) and the cell's exact
    // integer sums: xi (3); with +a also r (3), r x (q - xi) (3), r_a r_b
    // (xx yy zz xy xz yz)
    constexpr int NAT = ATA ? 15 : 3;
    __shared__ int s_xi[AT ? CAP : 1][3];
    __shared__ unsigned long long s_at[AT ? NC : 1][NAT];

    for (int c = tid; c < NC; c += kBlock) {
        s_cnt[c] = 0;
        s_owner[c] = -1;
        s_claim[c] = 0;
        if constexpr (BEADS) {
            s_bn[c] = 0;
            s_bfirst[c] = 0;
        }
        if constexpr (SOLID && BEADS)
            s_cut[c] = 0;
        if constexpr (AT) {
#pragma unroll
            for (int i = 0; i < NAT; ++i)
                s_at[c][i] = 0;
        }
    }
    tile_segments<W>(A, tc, s_seg, s_src, s_ocnt);
    tile_dests<W>(A, tc.t, s_dst);
    if (tid < 6)
        s_red[tid] = 0;
    if (tid == 0) {
        s_dred[0] = 0.0;
        s_anytie = 0;
        s_qn = 0;
    }
    bool sph_cut_tile = false;  // thread 0: the sphere cuts some cell of this tile
    if constexpr (SPHERE) {
        if (tid < 6)
            s_sph[tid] = 0.0;
        if (tid == 0) {
            const SphereDev *S = A.sphere;
            constexpr int T3[3] = {TX, TY, TZ};
            const double R2 = S->R * S->R;
            const double X[3] = {S->X[0], S->X[1], S->X[2]};
            const double Xh[3] = {S->X[0] + S->V[0] * A.h, S->X[1] + S->V[1] * A.h, S->X[2] + S->V[2] * A.h};
            sph_cut_tile = tile_dist2(A, tc, T3, X, 0.0) < R2;
            s_near = tile_dist2(A, tc, T3, Xh, kSolidReach + 0.5) < R2;
        }
    }
    if constexpr (WALLS) {
        if (tid == 0) {  // the wall layers' cells (z = 0, L_z - 1; bead pass), heights within reach of a wall
            const int Lz = g.L[2];
            sph_cut_tile = tc.oz == 0 || tc.oz + TZ >= Lz;
            s_near = tc.oz < (int)kSolidReach + 2 || tc.oz + TZ > Lz - 3 - (int)kSolidReach;
        }
    }
    __syncthreads();
    tile_admit<PASS>(A, s_seg, CAP, tc.t, D, &s_defer, sph_cut_tile);
    if constexpr (BEADS)
        tile_beads(A, tc.t, s_bfirst, s_bn);
    __syncthreads();
    if (s_defer)
        return;
    const int total = min(s_seg[27], CAP);

    uint32_t rec[K][W];
    load_tile<K, W>(s_seg, s_src, total, rec);

    // ---- while the records are on their way, every cell's axis and first
    //      two Gamma candidates (neither depends on the cell's particles):
    //      the normals by the cell's thread, the rest by a thread the cell
    //      sums will leave idle
    if (!AT && !A.inj_axis && (tid < NC || tid >= kBlock - NC)) {
        const int c = tid < NC ? tid : tid - (kBlock - NC);
        const int cx = tc.ox + c % TX, cy = tc.oy + (c / TX) % TY, cz = tc.oz + c / (TX * TY);
        if (cx < g.L[0] && cy < g.L[1] && cz < g.L[2]) {
            const long long gid = cx + (long long)g.L[0] * (cy + (long long)g.L[1] * cz);
            CellDrawsF32 &dr = *reinterpret_cast<CellDrawsF32 *>(&s_map[c][0]);
            if (tid < NC) {
                cell_normals_f32(A.seed, A.step, (uint64_t)gid, dr.x0, dr.x1);
            } else {
                double ax[3];
                uint32_t uw0, uw1;
                cell_axis_f32(A.seed, A.step, (uint64_t)gid, ax, uw0, uw1);
                dr.ax[0] = ax[0];
                dr.ax[1] = ax[1];
                dr.ax[2] = ax[2];
                dr.uw0 = uw0;
                dr.uw1 = uw1;
            }
        }
    }

    // ---- pass 1: rank in the cell, then the velocity to its slot
    uint32_t lcp[(K + 3) / 4];  // local cells, 8 bits each
    uint32_t slp[(K + 1) / 2];  // rank, then slot, 16 bits each
#pragma unroll
    for (int i = 0; i < (K + 3) / 4; ++i)
        lcp[i] = 0;
#pragma unroll
    for (int i = 0; i < (K + 1) / 2; ++i)
        slp[i] = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (tid + k * kBlock < total) {
            const int lc = local_cell<TX, TY, TZ>(A, tc, rec[k][6]);
            lcp[k >> 2] |= (uint32_t)lc << (8 * (k & 3));
            slp[k >> 1] |= (uint32_t)atomicAdd(&s_cnt[lc], 1) << (16 * (k & 1));
        }
    }
    __syncthreads();
    if (tid < 64)
        wave0_exclusive_scan<NC>(s_cnt, s_start);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (tid + k * kBlock < total) {
            const int lc = (lcp[k >> 2] >> (8 * (k & 3))) & 255;
            const int slot = s_start[lc] + (int)((slp[k >> 1] >> (16 * (k & 1))) & 0xffffu);
            slp[k >> 1] = (slp[k >> 1] & ~(0xffffu << (16 * (k & 1)))) | ((uint32_t)slot << (16 * (k & 1)));
            s_w[slot][0] = bitsf(rec[k][3]);
            s_w[slot][1] = bitsf(rec[k][4]);
            s_w[slot][2] = bitsf(rec[k][5]);
            if constexpr (AT) {
                const uint32_t key = rec[k][6];
                const long long gid = (long long)(key & 1023u) +
                                      (long long)g.L[0] * ((long long)((key >> 10) & 1023u) +
                                                           (long long)g.L[1] * (long long)(key >> 20));
                int xq[3];
                at_f32_fluid_xi<W, TRACK>(A, (uint64_t)gid, rec[k], xq);
                unsigned long long *S = s_at[lc];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    s_xi[slot][a] = xq[a];
                    atomicAdd(&S[a], (unsigned long long)(long long)xq[a]);
                }
                if constexpr (ATA) {
                    unsigned long long rq[3], dq[3];  // two's complement products: exact while sums fit 63 bits
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        rq[a] = (unsigned long long)at_f32_rq(rec[k][a]);
                        dq[a] = (unsigned long long)((long long)q32_of(bitsf(rec[k][3 + a])) - xq[a]);
                        atomicAdd(&S[3 + a], rq[a]);
                    }
                    atomicAdd(&S[6], rq[1] * dq[2] - rq[2] * dq[1]);
                    atomicAdd(&S[7], rq[2] * dq[0] - rq[0] * dq[2]);
                    atomicAdd(&S[8], rq[0] * dq[1] - rq[1] * dq[0]);
                    atomicAdd(&S[9], rq[0] * rq[0]);
                    atomicAdd(&S[10], rq[1] * rq[1]);
                    atomicAdd(&S[11], rq[2] * rq[2]);
                    atomicAdd(&S[12], rq[0] * rq[1]);
                    atomicAdd(&S[13], rq[0] * rq[2]);
                    atomicAdd(&S[14], rq[1] * rq[2]);
                }
            }
        }
    }
    __syncthreads();

    // ---- pass 2, per cell: exact sums and owner key (serial over the
    //      cell's slots), centre-of-mass velocity, MBS thermostat
    //      (mpcd.cpp:200-210), rotation (mpcd.cpp:269-295) as one affine map,
    //      all by the cell's thread (the draws are in LDS already)
    static_assert(2 * NC <= kBlock, "cell threads and the load-time draws' other half are disjoint");
    unsigned long long acc_e0 = 0, acc_e1 = 0;
    long long acc_m[3] = {0, 0, 0};
    double S1[3] = {0.0, 0.0, 0.0};  // kept by the cell's thread for pass 4
    double bmom[3] = {0.0, 0.0, 0.0};  // bead momentum of the cell before the collision
    unsigned long long lo = 0, hi = 0;  // sum of |q|^2 (q = v 2^21) split at bit 40
    bool cut = false;                   // the rigid sphere cuts the cell (bead pass)
    double dcen[3] = {0.0, 0.0, 0.0};   // then: the cell centre relative to the sphere's
    if (tid < NC) {
        const int c = tid;
        const int nfl = s_cnt[c];
        const int nbd = BEADS ? s_bn[c] : 0;
        const int p0 = s_start[c];
        if constexpr (SPHERE && BEADS) {
            const int cx = tc.ox + c % TX, cy = tc.oy + (c / TX) % TY, cz = tc.oz + c / (TX * TY);
            if (cx < g.L[0] && cy < g.L[1] && cz < g.L[2])
                cut = sphere_cuts_cell(A, cx, cy, cz, dcen);
            s_cut[c] = cut;
        }
        if constexpr (WALLS && BEADS) {
            // the bottom and top layers' cells below rho particles take the
            // wall's virtual mass (mpcd.cpp:175-198; the cell test as
            // written there)
            const int cx = tc.ox + c % TX, cy = tc.oy + (c / TX) % TY, cz = tc.oz + c / (TX * TY);
            const long long layer = (long long)g.L[0] * g.L[1];
            const long long gid = cx + (long long)g.L[0] * (cy + (long long)g.L[1] * cz);
            if (cx < g.L[0] && cy < g.L[1] && cz < g.L[2] && nfl + nbd > 0)
                cut = (gid < layer || gid > layer * (g.L[2] - 1)) && nfl + nbd < A.rho;
            s_cut[c] = cut;
        }
        // Sums in FP64, exact: every velocity is a multiple of 2^-21 below
        // 512 in magnitude, so sums of them below 2^32 are exact; squares
        // are multiples of 2^-42, and a sum of non-negative such terms is
        // exact while it stays below 2^11 (checked below; a larger one is
        // summed again in integers).  Exact sums do not depend on the order
        // of the slots.
        double se = 0.0;
        unsigned long long kmax = 0;
        int owner = -1, nmax = 0;
#pragma unroll 2
        for (int p = p0; p < p0 + nfl; ++p) {
            const float w0 = s_w[p][0], w1 = s_w[p][1], w2 = s_w[p][2];
            const double d0 = w0, d1 = w1, d2 = w2;
            S1[0] += d0;
            S1[1] += d1;
            S1[2] += d2;
            se += fma(d0, d0, fma(d1, d1, d2 * d2));
            const unsigned long long key = key2w(fbits(w0), fbits(w1));
            if (owner < 0 || key > kmax) {
                kmax = key;
                owner = p;
                nmax = 1;
            } else if (key == kmax) {
                ++nmax;
            }
        }
        if (se < 2048.0) {
            const unsigned long long e = (unsigned long long)(se * 4398046511104.0);
            lo = e & ((1ull << 40) - 1);
            hi = e >> 40;
        } else {
            for (int p = p0; p < p0 + nfl; ++p) {
                const int q0 = q32_of(s_w[p][0]), q1 = q32_of(s_w[p][1]), q2 = q32_of(s_w[p][2]);
                const unsigned long long e = (unsigned long long)((long long)q0 * q0) +
                                             (unsigned long long)((long long)q1 * q1) +
                                             (unsigned long long)((long long)q2 * q2);
                lo += e & ((1ull << 40) - 1);
                hi += e >> 40;
            }
        }
        if (nmax > 1) {  // stage-0 tie: the candidates settle it in pass 4
            s_k1[c] = kmax;
            s_anytie = 1;
            owner = -1;
        }
        if ((!BEADS || nbd == 0) && !cut)
            s_owner[c] = owner;
    }
    if (tid < NC) {  // the same thread: no barrier
        const int c = tid;
        const int cx = tc.ox + c % TX, cy = tc.oy + (c / TX) % TY, cz = tc.oz + c / (TX * TY);
        const int nfl = s_cnt[c];
        const int nbd = BEADS ? s_bn[c] : 0;
        const int num = nfl + nbd;
        if (AT && num > 0 && cx < g.L[0] && cy < g.L[1] && cz < g.L[2]) {
            // MPC-AT(+a) (Noguchi, Kikuchi, Gompper, EPL 78 (2007) 10005) as
            // v' = b + xi + omega x r with b = u - <xi> - omega x C: fluid sums
            // exact, bead sums in FP64 in the beads' order
            const long long gid = cx + (long long)g.L[0] * (cy + (long long)g.L[1] * cz);
            double bmv[3] = {0.0, 0.0, 0.0}, bX[3] = {0.0, 0.0, 0.0}, bC[3] = {0.0, 0.0, 0.0};
            double bL[3] = {0.0, 0.0, 0.0}, bI[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, ble = 0.0;
            int bm = 0;
            if constexpr (BEADS) {
                for (int i = s_bfirst[c]; i < s_bfirst[c] + nbd; ++i) {
                    const double *v = A.b_vel4 + 4 * (long long)i;
                    const double m = v[3];
                    double xi[3];
                    at_f32_bead_xi(A, (uint64_t)gid, i, xi);
                    bmv[0] += v[0] * m;
                    bmv[1] += v[1] * m;
                    bmv[2] += v[2] * m;
                    bm += (int)m;
                    ble += m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
#pragma unroll
                    for (int a = 0; a < 3; ++a)
                        bX[a] += m * xi[a];
                    if constexpr (ATA) {
                        double r[3], d[3];
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            r[a] = local_coord(A.b_pos3[3 * (long long)A.b_idx[i] + a], A.a[a]);
                            d[a] = v[a] - xi[a];
                            bC[a] += m * r[a];
                        }
                        bL[0] += m * (r[1] * d[2] - r[2] * d[1]);
                        bL[1] += m * (r[2] * d[0] - r[0] * d[2]);
                        bL[2] += m * (r[0] * d[1] - r[1] * d[0]);
                        bI[0] += m * (r[1] * r[1] + r[2] * r[2]);
                        bI[1] += m * (r[0] * r[0] + r[2] * r[2]);
                        bI[2] += m * (r[0] * r[0] + r[1] * r[1]);
                        bI[3] -= m * (r[0] * r[1]);
                        bI[4] -= m * (r[0] * r[2]);
                        bI[5] -= m * (r[1] * r[2]);
                    }
                }
                if (nbd)
                    atomicAdd(&s_dred[0], ble);
                bmom[0] = bmv[0];
                bmom[1] = bmv[1];
                bmom[2] = bmv[2];
            }
            const double Mc = (double)(nfl + bm);
            const unsigned long long *S = s_at[c];
            double u[3], xm[3], om[3] = {0.0, 0.0, 0.0}, bp[3];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                u[a] = (S1[a] + bmv[a]) / Mc;
                xm[a] = ((double)(long long)S[a] * kQinv + bX[a]) / Mc;
            }
            if constexpr (ATA) {
                double C[3], Dm[3], L[3], I[6];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    C[a] = ((double)(long long)S[3 + a] * kRqinv + bC[a]) / Mc;
                    Dm[a] = (S1[a] + bmv[a]) - ((double)(long long)S[a] * kQinv + bX[a]);
                }
                L[0] = ((double)(long long)S[6] * kRQinv + bL[0]) - (C[1] * Dm[2] - C[2] * Dm[1]);
                L[1] = ((double)(long long)S[7] * kRQinv + bL[1]) - (C[2] * Dm[0] - C[0] * Dm[2]);
                L[2] = ((double)(long long)S[8] * kRQinv + bL[2]) - (C[0] * Dm[1] - C[1] * Dm[0]);
                I[0] = ((double)(long long)(S[10] + S[11]) * kRRinv + bI[0]) - Mc * (C[1] * C[1] + C[2] * C[2]);
                I[1] = ((double)(long long)(S[9] + S[11]) * kRRinv + bI[1]) - Mc * (C[0] * C[0] + C[2] * C[2]);
                I[2] = ((double)(long long)(S[9] + S[10]) * kRRinv + bI[2]) - Mc * (C[0] * C[0] + C[1] * C[1]);
                I[3] = (bI[3] - (double)(long long)S[12] * kRRinv) + Mc * (C[0] * C[1]);
                I[4] = (bI[4] - (double)(long long)S[13] * kRRinv) + Mc * (C[0] * C[2]);
                I[5] = (bI[5] - (double)(long long)S[14] * kRRinv) + Mc * (C[1] * C[2]);
                at_omega(I, L, om);
                bp[0] = (u[0] - xm[0]) - (om[1] * C[2] - om[2] * C[1]);
                bp[1] = (u[1] - xm[1]) - (om[2] * C[0] - om[0] * C[2]);
                bp[2] = (u[2] - xm[2]) - (om[0] * C[1] - om[1] * C[0]);
            } else {
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    bp[a] = u[a] - xm[a];
            }
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                s_map[c][a] = (float)bp[a];
                s_map[c][3 + a] = (float)om[a];
            }
            if (BEADS && nbd) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    s_map64[c][a] = bp[a];
                    s_map64[c][3 + a] = om[a];
                }
            }
            acc_e0 = (hi << 40) + lo;
            acc_m[0] = (long long)(S1[0] * kQ);
            acc_m[1] = (long long)(S1[1] * kQ);
            acc_m[2] = (long long)(S1[2] * kQ);
        } else if (!AT && num > 0 && cx < g.L[0] && cy < g.L[1] && cz < g.L[2]) {
            double bx = 0.0, by = 0.0, bz = 0.0, ble = 0.0, vle = 0.0;
            int bm = 0, nv = 0;
            if constexpr (BEADS) {
                for (int i = s_bfirst[c]; i < s_bfirst[c] + nbd; ++i) {
                    const double *v = A.b_vel4 + 4 * (long long)i;
                    const double m = v[3];
                    bx += v[0] * m;
                    by += v[1] * m;
                    bz += v[2] * m;
                    bm += (int)m;
                    ble += m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                }
                if (nbd)
                    atomicAdd(&s_dred[0], ble);
                bmom[0] = bx;
                bmom[1] = by;
                bmom[2] = bz;
            }
            if constexpr (SPHERE && BEADS) {
                // of rho candidate points in a cell the sphere cuts, those
                // inside are virtual particles moving with it, plus noise
                if (cut) {
                    for (int q = 0; q < A.rho; ++q) {
                        const SphereVirtual sv = sphere_virtual(A, cx + (long long)g.L[0] * (cy + (long long)g.L[1] * cz),
                                                                dcen[0], dcen[1], dcen[2], q);
                        if (!sv.in)
                            continue;
                        bx += sv.v0;
                        by += sv.v1;
                        bz += sv.v2;
                        vle += sv.v0 * sv.v0 + sv.v1 * sv.v1 + sv.v2 * sv.v2;
                        ++nv;
                    }
                }
            }
            if constexpr (WALLS && BEADS) {
                if (cut) {  // rho - num virtual particles of momentum sqrt((rho - num) kT) N(0,1) per axis
                    const long long gidw = cx + (long long)g.L[0] * (cy + (long long)g.L[1] * cz);
                    const int mass_diff = A.rho - num;
                    double nr[3];
                    if (A.inj_virt) {
                        nr[0] = A.inj_virt[3 * gidw];
                        nr[1] = A.inj_virt[3 * gidw + 1];
                        nr[2] = A.inj_virt[3 * gidw + 2];
                    } else {
                        virt_normals(A.seed, A.step, (uint64_t)gidw, nr);
                    }
                    const double sv = sqrt(mass_diff * A.kT);
                    const double v0 = sv * nr[0], v1 = sv * nr[1], v2 = sv * nr[2];
                    bx += v0;
                    by += v1;
                    bz += v2;
                    vle += (v0 * v0 + v1 * v1 + v2 * v2) / mass_diff;
                    nv += mass_diff;
                }
            }
            const int masse = nfl + bm + nv;
            const double u[3] = {(S1[0] + bx) / masse, (S1[1] + by) / masse, (S1[2] + bz) / masse};
            const double s2 = (double)hi * 1099511627776.0 + (double)lo;
            const double le = s2 * (kQinv * kQinv) + ble + vle;
            const double rel = le - (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) * masse;
            const long long gid = cx + (long long)g.L[0] * (cy + (long long)g.L[1] * cz);
            double ax[3], G;
            if (A.inj_axis) {
                ax[0] = A.inj_axis[3 * gid];
                ax[1] = A.inj_axis[3 * gid + 1];
                ax[2] = A.inj_axis[3 * gid + 2];
                G = A.inj_gamma[gid];
            } else {
                // the candidates drawn during the load become the Gamma variate
                const CellDrawsF32 &dr = *reinterpret_cast<const CellDrawsF32 *>(&s_map[c][0]);
                G = A.thermostat && num + nv > 1
                        ? philox_gamma_f32(A.seed, A.step, (uint64_t)gid, num + nv, dr.x0, dr.uw0, dr.x1, dr.uw1)
                        : 0.0;
                ax[0] = dr.ax[0];
                ax[1] = dr.ax[1];
                ax[2] = dr.ax[2];
            }
            double R[9];
            const double sc = scale_rotation(A, G, ax, num + nv, rel, D, R);
            double M[9];
#pragma unroll
            for (int i = 0; i < 9; ++i)
                M[i] = R[i] * sc;
            double b[3];
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                double t = M[3 * i + 2] * u[2];
                t = fma(M[3 * i + 1], u[1], t);
                t = fma(M[3 * i], u[0], t);
                b[i] = u[i] - t;
                s_map[c][9 + i] = (float)b[i];
            }
#pragma unroll
            for (int i = 0; i < 9; ++i)
                s_map[c][i] = (float)M[i];
            if (BEADS && nbd) {
#pragma unroll
                for (int i = 0; i < 9; ++i)
                    s_map64[c][i] = M[i];
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    s_map64[c][9 + i] = b[i];
            }
            if constexpr (SPHERE && BEADS) {
                if (nv) {  // the virtual particles' angular impulse goes to the sphere
                    double T[3] = {0.0, 0.0, 0.0};
                    for (int q = 0; q < A.rho; ++q) {
                        const SphereVirtual sv = sphere_virtual(A, gid, dcen[0], dcen[1], dcen[2], q);
                        if (!sv.in)
                            continue;
                        const double v[3] = {sv.v0, sv.v1, sv.v2};
                        double Jq[3];
#pragma unroll
                        for (int a = 0; a < 3; ++a)
                            Jq[a] = fma(M[3 * a], v[0], fma(M[3 * a + 1], v[1], fma(M[3 * a + 2], v[2], b[a]))) - v[a];
                        T[0] += sv.d1 * Jq[2] - sv.d2 * Jq[1];
                        T[1] += sv.d2 * Jq[0] - sv.d0 * Jq[2];
                        T[2] += sv.d0 * Jq[1] - sv.d1 * Jq[0];
                    }
#pragma unroll
                    for (int a = 0; a < 3; ++a)
                        atomicAdd(&s_sph[3 + a], T[a]);
                }
            }
            acc_e0 = (hi << 40) + lo;
            acc_m[0] = (long long)(S1[0] * kQ);
            acc_m[1] = (long long)(S1[1] * kQ);
            acc_m[2] = (long long)(S1[2] * kQ);
        }
    }
    __syncthreads();

    // ---- pass 3: new velocities into the slots; every particle streams with
    //      its new velocity and is written, except the stage-0 candidates of
    //      tied cells, which are queued, and the owner of every other fluid
    //      cell, whose streamed record waits in LDS for the residual
    uint32_t dmask = 0;
    bool bad_v = false;
    const bool interior = tile_interior(g, tc);
    long long sph_dm[3] = {0, 0, 0}, sph_de = 0;  // the fluid's momentum and |q|^2 changed by bounces off a solid
#pragma unroll
    for (int k = 0; k < K; ++k) {
        int dest = -1;
        uint32_t r[W];
        bool cand = false, mine = false;
        int lc = 0;
        float w[3] = {0.0f, 0.0f, 0.0f};
        if (tid + k * kBlock < total) {
            lc = (lcp[k >> 2] >> (8 * (k & 3))) & 255;
            const int slot = (int)((slp[k >> 1] >> (16 * (k & 1))) & 0xffffu);
            const float *M = s_map[lc];
            // the old velocity from its slot: the record's velocity words are
            // not needed in registers after pass 1
            const float v0 = s_w[slot][0], v1 = s_w[slot][1], v2 = s_w[slot][2];
            int qn[3];
            if constexpr (AT) {
                float vn[3], rf[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    vn[i] = __fadd_rn(M[i], (float)s_xi[slot][i] * (1.0f / 2097152.0f));
                    rf[i] = (float)at_f32_rq(rec[k][i]) * (1.0f / 16777216.0f);
                }
                if constexpr (ATA) {
#pragma unroll
                    for (int i = 0; i < 3; ++i) {
                        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
                        vn[i] = fmaf(M[3 + i1], rf[i2], vn[i]);
                        vn[i] = fmaf(-M[3 + i2], rf[i1], vn[i]);
                    }
                }
#pragma unroll
                for (int i = 0; i < 3; ++i)
                    w[i] = quant_fq(vn[i], qn[i]);
            } else {
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    const float vn = fmaf(M[3 * i], v0, fmaf(M[3 * i + 1], v1, fmaf(M[3 * i + 2], v2, M[9 + i])));
                    w[i] = quant_fq(vn, qn[i]);
                }
            }
            const int own = s_owner[lc];
            mine = own == slot;
            cand = own < 0 && (!BEADS || s_bn[lc] == 0) && !(SOLID && BEADS && s_cut[lc]) &&
                   key2w(fbits(v0), fbits(v1)) == s_k1[lc];
            s_w[slot][0] = w[0];
            s_w[slot][1] = w[1];
            s_w[slot][2] = w[2];
            if (cand) {
                dmask |= 1u << k;
            } else {
                dest = interior ? stream_rebin_interior<TX, TY, TZ, W>(A, tc, lc, rec[k], w, r, bad_v)
                                : stream_rebin<TX, TY, TZ, W>(A, tc, rec[k], w, r, bad_v);
                if constexpr (SOLID) {
                    if (s_near || solid_fast(A, w))
                        dest = solid_bounce_f32<OBJ, TX, TY, TZ, W>(A, tc, rec[k], w, r, dest, s_sph, sph_dm, sph_de,
                                                                    bad_v);
                }
                if (mine) {
#pragma unroll
                    for (int i = 0; i < W; ++i)
                        s_orec[lc][i] = r[i];
                    s_orec[lc][W] = (uint32_t)dest;
                    dest = -1;
                }
                if constexpr (TRACK) {
                    if (A.dbg_cell) {
                        const uint32_t key = rec[k][6];
                        A.dbg_cell[rec[k][7]] =
                            (int)(key & 1023u) + g.L[0] * ((int)((key >> 10) & 1023u) + g.L[1] * (int)(key >> 20));
                        A.dbg_rank[rec[k][7]] = -1;
                    }
                }
            }
        }
        const int qs = wave_bump(&s_qn, cand);
        if (cand && qs < QCAP) {
#pragma unroll
            for (int i = 0; i < W; ++i)
                s_q[qs][i] = i >= 3 && i < 6 ? fbits(w[i - 3]) : rec[k][i];
            s_q[qs][W] = (uint32_t)lc;
        }
        place_record<W>(A, s_dst, s_ocnt, dest, r);
    }
    __syncthreads();

    // ---- pass 4, per cell: residual sum (q - q') over the cell's slots and
    //      the sum of |q|^2 after the collision (exact, as the sums of pass
    //      2); the owner of a fluid-only cell keeps its new velocity plus the
    //      residual in its streamed record, which is written now
    if (tid < NC) {
        const int c = tid;
        const int p0 = s_start[c], nfl = s_cnt[c];
        double Sn[3] = {0.0, 0.0, 0.0};
        double se1 = 0.0;
#pragma unroll 2
        for (int p = p0; p < p0 + nfl; ++p) {
            const double d0 = s_w[p][0], d1 = s_w[p][1], d2 = s_w[p][2];
            Sn[0] += d0;
            Sn[1] += d1;
            Sn[2] += d2;
            se1 += fma(d0, d0, fma(d1, d1, d2 * d2));
        }
        if (se1 < 2048.0) {
            acc_e1 = (unsigned long long)(se1 * 4398046511104.0);
        } else {
            unsigned long long lo1 = 0, hi1 = 0;
            for (int p = p0; p < p0 + nfl; ++p) {
                const int q0 = q32_of(s_w[p][0]), q1 = q32_of(s_w[p][1]), q2 = q32_of(s_w[p][2]);
                const unsigned long long e = (unsigned long long)((long long)q0 * q0) +
                                             (unsigned long long)((long long)q1 * q1) +
                                             (unsigned long long)((long long)q2 * q2);
                lo1 += e & ((1ull << 40) - 1);
                hi1 += e >> 40;
            }
            acc_e1 = (hi1 << 40) + lo1;
        }
        int dq[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            dq[a] = (int)((S1[a] - Sn[a]) * kQ);
            s_dq[c][a] = dq[a];
        }
        if (BEADS && (s_bn[c] > 0 || cut)) {
#pragma unroll
            for (int a = 0; a < 3; ++a)
                acc_m[a] -= dq[a];
        }
        if constexpr (SPHERE && BEADS) {
            if (cut && s_bn[c] == 0) {  // the fluid's residual goes to the sphere
#pragma unroll
                for (int a = 0; a < 3; ++a)
                    atomicAdd(&s_sph[a], (double)dq[a] * kQinv);
            }
        }
        // the owner of a fluid-only cell keeps its new velocity plus the
        // residual in its streamed record, which is written now
        const int os = s_owner[c];
        int dest = -1;
        uint32_t r[W];
        if (os >= 0) {
#pragma unroll
            for (int i = 0; i < W; ++i)
                r[i] = s_orec[c][i];
            dest = (int)s_orec[c][W];
            const float w[3] = {bitsf(r[3]), bitsf(r[4]), bitsf(r[5])};
            float wf[3];
            long long de = 0;
            owner_take_residual(w, dq, wf, de, acc_m);
            acc_e1 += (unsigned long long)de;
#pragma unroll
            for (int a = 0; a < 3; ++a)
                r[3 + a] = fbits(wf[a]);
        }
        place_record<W>(A, s_dst, s_ocnt, dest, r);
    }
    __syncthreads();

    // ---- the deferred records: the owner of each tied cell is settled by
    //      the stage keys of the new records, then the id, then (records now
    //      identical) the first to claim; owners take the residual; all are
    //      written.  Normally they sit in the LDS queue, one per thread; a
    //      tile with more than QCAP of them uses the records in registers.
    if (s_anytie) {  // the map is no longer needed: its words take the stage-key maxima
        for (int i = tid; i < (NSTAGE - 1) * NC; i += kBlock)
            s_kst[i / NC][i % NC] = 0;
        __syncthreads();
    }
    const int qn = s_qn;
    if (qn <= QCAP) {
        const bool live = tid < qn;
        uint32_t q[W];
        int lc = 0;
        if (live) {
#pragma unroll
            for (int i = 0; i < W; ++i)
                q[i] = s_q[tid][i];
            lc = (int)s_q[tid][W];
        }
        bool is_owner = live;
        if (s_anytie) {
            const bool tied = live && s_owner[lc] < 0;
#pragma unroll
            for (int st = 1; st < NSTAGE; ++st) {
                if (tied && is_owner)
                    atomicMax(&s_kst[st - 1][lc], owner_stage_key<W>(st, q));
                __syncthreads();
                if (tied && is_owner && owner_stage_key<W>(st, q) != s_kst[st - 1][lc])
                    is_owner = false;
            }
            if (tied && is_owner && atomicAdd(&s_claim[lc], 1) != 0)
                is_owner = false;
        }
        if ((tid & ~63) < qn) {
            int dest = -1;
            uint32_t r[W];
            if (live)
                if constexpr (SOLID)
                    dest = finish_deferred_obj<OBJ, TX, TY, TZ, W, TRACK>(A, tc, q, s_dq[lc], is_owner, acc_e1, acc_m,
                                                                          r, bad_v, s_near, s_sph, sph_dm, sph_de);
                else
                    dest = finish_deferred<TX, TY, TZ, W, TRACK>(A, tc, q, s_dq[lc], is_owner, acc_e1, acc_m, r, bad_v);
            place_record<W>(A, s_dst, s_ocnt, dest, r);
        }
    } else {
        // the new velocities of the deferred records come back from their slots
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if ((dmask >> k) & 1u) {
                const int slot = (int)((slp[k >> 1] >> (16 * (k & 1))) & 0xffffu);
                rec[k][3] = fbits(s_w[slot][0]);
                rec[k][4] = fbits(s_w[slot][1]);
                rec[k][5] = fbits(s_w[slot][2]);
            }
        }
        uint32_t owner = dmask;
        if (s_anytie) {
#pragma unroll
            for (int st = 1; st < NSTAGE; ++st) {
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    if ((owner >> k) & 1u) {
                        const int lc = (lcp[k >> 2] >> (8 * (k & 3))) & 255;
                        if (s_owner[lc] < 0)
                            atomicMax(&s_kst[st - 1][lc], owner_stage_key<W>(st, rec[k]));
                    }
                }
                __syncthreads();
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    if ((owner >> k) & 1u) {
                        const int lc = (lcp[k >> 2] >> (8 * (k & 3))) & 255;
                        if (s_owner[lc] < 0 && owner_stage_key<W>(st, rec[k]) != s_kst[st - 1][lc])
                            owner &= ~(1u << k);
                    }
                }
            }
#pragma unroll
            for (int k = 0; k < K; ++k) {
                if ((owner >> k) & 1u) {
                    const int lc = (lcp[k >> 2] >> (8 * (k & 3))) & 255;
                    if (s_owner[lc] < 0 && atomicAdd(&s_claim[lc], 1) != 0)
                        owner &= ~(1u << k);
                }
            }
        }
#pragma unroll
        for (int k = 0; k < K; ++k) {
            if (__ballot((dmask >> k) & 1u) == 0)
                continue;
            int dest = -1;
            uint32_t r[W];
            if ((dmask >> k) & 1u) {
                const int lc = (lcp[k >> 2] >> (8 * (k & 3))) & 255;
                if constexpr (SOLID)
                    dest = finish_deferred_obj<OBJ, TX, TY, TZ, W, TRACK>(A, tc, rec[k], s_dq[lc], (owner >> k) & 1u,
                                                                          acc_e1, acc_m, r, bad_v, s_near, s_sph,
                                                                          sph_dm, sph_de);
                else
                    dest = finish_deferred<TX, TY, TZ, W, TRACK>(A, tc, rec[k], s_dq[lc], (owner >> k) & 1u, acc_e1,
                                                                 acc_m, r, bad_v);
            }
            place_record<W>(A, s_dst, s_ocnt, dest, r);
        }
    }

    // ---- beads: the FP64 map; then the first bead of the cell takes the
    //      cell's momentum balance, so fluid (exact integers) + beads
    //      conserve momentum to FP64 roundoff
    if constexpr (BEADS) {
        if (tid < NC && s_bn[tid] > 0) {
            const int c = tid;
            const double *M = s_map64[c];
            const int i0 = s_bfirst[c], i1 = s_bfirst[c] + s_bn[c];
            const long long gid = (tc.ox + c % TX) + (long long)g.L[0] * ((tc.oy + (c / TX) % TY) +
                                                                          (long long)g.L[1] * (tc.oz + c / (TX * TY)));
            double pb[3] = {0.0, 0.0, 0.0}, n0[3] = {0.0, 0.0, 0.0};
            for (int i = i0; i < i1; ++i) {
                const double *v = A.b_vel4 + 4 * (long long)i;
                const double m = v[3];
                double n[3];
                if constexpr (AT) {
                    double xi[3];
                    at_f32_bead_xi(A, (uint64_t)gid, i, xi);
                    if constexpr (ATA) {
                        double r[3];
#pragma unroll
                        for (int a = 0; a < 3; ++a)
                            r[a] = local_coord(A.b_pos3[3 * (long long)A.b_idx[i] + a], A.a[a]);
#pragma unroll
                        for (int a = 0; a < 3; ++a) {
                            const int a1 = (a + 1) % 3, a2 = (a + 2) % 3;
                            n[a] = (M[a] + xi[a]) + (M[3 + a1] * r[a2] - M[3 + a2] * r[a1]);
                        }
                    } else {
#pragma unroll
                        for (int a = 0; a < 3; ++a)
                            n[a] = M[a] + xi[a];
                    }
#pragma unroll
                    for (int a = 0; a < 3; ++a)
                        pb[a] += n[a] * m;
                } else {
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        n[a] = fma(M[3 * a], v[0], fma(M[3 * a + 1], v[1], fma(M[3 * a + 2], v[2], M[9 + a])));
                        pb[a] += n[a] * m;
                    }
                }
                double *o = A.b_out4 + 4 * (long long)A.b_idx[i];
                o[0] = n[0];
                o[1] = n[1];
                o[2] = n[2];
                o[3] = m;
                if (i == i0) {
                    n0[0] = n[0];
                    n0[1] = n[1];
                    n0[2] = n[2];
                }
            }
            {
                const double m0 = A.b_vel4[4 * (long long)i0 + 3];
                double *o = A.b_out4 + 4 * (long long)A.b_idx[i0];
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const double before = S1[a] + bmom[a];
                    const double after = (S1[a] - (double)s_dq[c][a] * kQinv) + pb[a];
                    o[a] = n0[a] + (before - after) / m0;
                }
            }
            double blen = 0.0;
            for (int i = i0; i < i1; ++i) {
                const double *o = A.b_out4 + 4 * (long long)A.b_idx[i];
                blen += o[3] * (o[0] * o[0] + o[1] * o[1] + o[2] * o[2]);
            }
            atomicAdd(&s_dred[0], -blen);
        }
    }
    if (bad_v)
        atomicOr(A.flags, kFlagVelocity);
    if constexpr (SOLID) {  // the bounces' exact changes of the fluid's momentum and |q|^2
#pragma unroll
        for (int a = 0; a < 3; ++a)
            if (sph_dm[a])
                atomicAdd(&s_red[2 + a], (unsigned long long)sph_dm[a]);
        if (sph_de)
            atomicAdd(&s_red[5], (unsigned long long)sph_de);
    }

    // ---- tile diagnostics
    acc_e0 = wave_sum_u64(acc_e0);
    acc_e1 = wave_sum_u64(acc_e1);
#pragma unroll
    for (int a = 0; a < 3; ++a)
        acc_m[a] = (long long)wave_sum_u64((unsigned long long)acc_m[a]);
    if (lane_id() == 0) {
        atomicAdd(&s_red[0], acc_e0);
        atomicAdd(&s_red[1], acc_e1);
        atomicAdd(&s_red[2], (unsigned long long)acc_m[0]);
        atomicAdd(&s_red[3], (unsigned long long)acc_m[1]);
        atomicAdd(&s_red[4], (unsigned long long)acc_m[2]);
    }
    __syncthreads();
    write_counts(A, tc.t, s_ocnt);
    if (tid == 0) {
        atomicAdd((unsigned long long *)&D->q_diss, s_red[0] - s_red[1]);
        atomicAdd((unsigned long long *)&D->q_mom[0], s_red[2]);
        atomicAdd((unsigned long long *)&D->q_mom[1], s_red[3]);
        atomicAdd((unsigned long long *)&D->q_mom[2], s_red[4]);
        atomicAdd(&D->e2, (double)(s_red[1] + s_red[5]) * (kQinv * kQinv));
        atomicAdd((unsigned long long *)&D->n, (unsigned long long)total);
        if (BEADS && s_dred[0] != 0.0)
            atomicAdd(&D->diss, s_dred[0]);
        if constexpr (SPHERE) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                atomicAdd(&A.sphere->dP[k], s_sph[k]);
                atomicAdd(&A.sphere->dL[k], s_sph[3 + k]);
            }
        }
    }
}

// ---------------------------------------------------------------- MPC-AT
// Cell-local coordinate of a position along one axis: (x + a) - floor(x + a),
// the offset inside its shifted cell, rounded to FP32 (the lever arm of the
// angular-momentum correction).
__device__ __forceinline__ double local_coord(double x, double a)
{
    const double t = x + a;
    return (double)(float)(t - floor(t));
}

// omega = I^-1 L for a cell's inertia tensor about its centre of mass
// (I: xx yy zz xy xz yz, off-diagonals signed): adjugate over determinant;
// for (nearly) collinear particles I = lambda (1 - e e^T), lambda = tr I / 2,
// and L is normal to e, so omega = L / lambda; zero for a point cell.
__device__ __forceinline__ void at_omega(const double I[6], const double L[3], double w[3])
{
    const double xx = I[0], yy = I[1], zz = I[2], xy = I[3], xz = I[4], yz = I[5];
    const double a00 = yy * zz - yz * yz, a01 = xz * yz - xy * zz, a02 = xy * yz - xz * yy;
    const double a11 = xx * zz - xz * xz, a12 = xy * xz - xx * yz, a22 = xx * yy - xy * xy;
    const double det = xx * a00 + xy * a01 + xz * a02;
    const double lam = (xx + yy + zz) * 0.5;
    if (det > 1e-12 * (lam * lam * lam)) {
        w[0] = (a00 * L[0] + a01 * L[1] + a02 * L[2]) / det;
        w[1] = (a01 * L[0] + a11 * L[1] + a12 * L[2]) / det;
        w[2] = (a02 * L[0] + a12 * L[1] + a22 * L[2]) / det;
    } else if (lam > 0) {
        w[0] = L[0] / lam;
        w[1] = L[1] / lam;
        w[2] = L[2] / lam;
    } else {
        w[0] = w[1] = w[2] = 0.0;
    }
}

// The three normals of the MPC-AT random velocity of the particle of rank r
// in its cell (FP64 form): Philox4x32-10 of counter {cell, low and high word
// of the step, r}, key {low word of the seed ^ 'MPAT', high word of the
// seed}; the cosine and sine branches of Box-Muller on words 0 and 1, the
// cosine branch on words 2 and 3, in FP64 on 32-bit uniforms.
constexpr uint64_t kTagMpat = 0x4d504154ull;

__device__ __forceinline__ void ran_normals(uint64_t seed, uint64_t step, uint64_t cell, uint64_t r, double n[3])
{
    const U32x4 w = philox4x32((uint32_t)cell, (uint32_t)step, (uint32_t)(step >> 32), (uint32_t)r,
                               (uint32_t)seed ^ (uint32_t)kTagMpat, (uint32_t)(seed >> 32));
    const double r0 = sqrt(-2.0 * log(u32_open01(w.v[0])));
    double s1, c1;
    sincospi(2.0 * u32_open01(w.v[1]), &s1, &c1);
    n[0] = r0 * c1;
    n[1] = r0 * s1;
    n[2] = sqrt(-2.0 * log(u32_open01(w.v[2]))) * cospi(2.0 * u32_open01(w.v[3]));
}

// ===================================================================== FP64
// Record (12 words, 14 with ids): x y z vx vy vz as doubles (positions
// unwrapped, as the reference keeps them), [id, pad].
template <int TX, int TY, int TZ, int K, bool TRACK, bool BEADS, int RULE, bool SPHERE, int PASS>
__device__ __forceinline__ void step_f64_tile(const StepArgs &A, int tile)
{
    constexpr bool AT = RULE != kRuleSrd;
    constexpr int NC = TX * TY * TZ;
    static_assert(NC <= kBlock, "one thread per cell");
    constexpr int W = TRACK ? 14 : 12;
    constexpr int CAP = K * kBlock;
    static_assert(CAP < 65536, "16-bit slots");
    constexpr int NB = BEADS ? NC : 1;
    const Geom &g = A.g;
    const int tid = threadIdx.x;
    const TileCoord tc = tile_coord(g, tile);
    DiagSlot *D = &A.diag->slot[tc.t & (kDiagSlots - 1)];

    __shared__ int s_seg[28];
    __shared__ const uint32_t *s_src[27];
    __shared__ int s_ocnt[27];
    __shared__ DestTable s_dst;
    __shared__ int s_cnt[NC];
    __shared__ int s_start[NC];
    __shared__ double s_v[CAP][3];
    __shared__ int s_id[TRACK ? CAP : 1];
    __shared__ unsigned short s_sorted[CAP];
    // each particle's order key at its slot in the cell's range while the
    // ranks are counted; then (SRD) the cells' fluid sums, one row per
    // component: Σvx, Σvy, Σvz, Σ|v|^2 (keys: ids, or the bits of vx)
    using OrderKey = std::conditional_t<TRACK, uint32_t, unsigned long long>;
    static_assert(AT || sizeof(OrderKey) * CAP >= sizeof(double) * 4 * NC, "the fluid sums fit the key array");
    __shared__ __align__(8) OrderKey s_key[CAP];
    __shared__ int s_bfirst[NB];
    __shared__ int s_bn[NB];
    __shared__ double s_dred[5];  // diss, fluid e after, fluid momentum xyz
    __shared__ double s_sph[6];   // rigid sphere: impulse and its moment from this tile
    // MPC-AT: each particle's random velocity, cell-local coordinate, rank
    __shared__ double s_xi[AT ? CAP : 1][3];
    __shared__ float s_rl[AT ? CAP : 1][3];
    __shared__ unsigned short s_rank[AT ? CAP : 1];

    for (int c = tid; c < NC; c += kBlock) {
        s_cnt[c] = 0;
        if constexpr (BEADS) {
            s_bn[c] = 0;
            s_bfirst[c] = 0;
        }
    }
    tile_segments<W>(A, tc, s_seg, s_src, s_ocnt);
    tile_dests<W>(A, tc.t, s_dst);
    if (tid < 5)
        s_dred[tid] = 0.0;
    if (tid < 6)
        s_sph[tid] = 0.0;
    __shared__ int s_defer;
    __syncthreads();
    tile_admit<PASS>(A, s_seg, CAP, tc.t, D, &s_defer);
    if constexpr (BEADS)
        tile_beads(A, tc.t, s_bfirst, s_bn);
    __syncthreads();
    if (s_defer)
        return;
    const int total = min(s_seg[27], CAP);

    // ---- 1. load; keep positions (+ id), velocities to LDS
    constexpr int PW = TRACK ? 7 : 6;
    uint32_t pk[K][PW];
    int slot_rank[K];
    {
        uint32_t rec[K][W];
        load_tile<K, W>(s_seg, s_src, total, rec);
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int j = tid + k * kBlock;
            if (j < total) {
                const double x = djoin(rec[k][0], rec[k][1]);
                const double y = djoin(rec[k][2], rec[k][3]);
                const double z = djoin(rec[k][4], rec[k][5]);
                int lx = box_axis_near(x + A.a[0], g.L[0], g.invL[0]) - tc.ox;
                int ly = box_axis_near(y + A.a[1], g.L[1], g.invL[1]) - tc.oy;
                int lz = box_axis_near(z + A.a[2], g.L[2], g.invL[2]) - tc.oz;
                if (lx < 0 || lx >= TX || ly < 0 || ly >= TY || lz < 0 || lz >= TZ) {
                    atomicOr(A.flags, kFlagMisbinned);
                    lx = min(max(lx, 0), TX - 1);
                    ly = min(max(ly, 0), TY - 1);
                    lz = min(max(lz, 0), TZ - 1);
                }
                const int lc = lx + TX * (ly + TY * lz);
                slot_rank[k] = (lc << 16) | atomicAdd(&s_cnt[lc], 1);
                s_v[j][0] = djoin(rec[k][6], rec[k][7]);
                s_v[j][1] = djoin(rec[k][8], rec[k][9]);
                s_v[j][2] = djoin(rec[k][10], rec[k][11]);
#pragma unroll
                for (int w = 0; w < 6; ++w)
                    pk[k][w] = rec[k][w];
                if constexpr (TRACK) {
                    pk[k][6] = rec[k][12];
                    s_id[j] = (int)rec[k][12];
                }
            }
        }
    }
    __syncthreads();
    if (tid < 64)
        wave0_exclusive_scan<NC>(s_cnt, s_start);
    __syncthreads();
    // ---- the stable order in each cell (the reference's insertion order):
    //      particle id (upload order) with ids, else the bits of vx, ties by
    //      arrival; beads follow in bead index order.  The keys go to their
    //      slots in the cell's range (arrival order), then each particle
    //      counts the keys before its own
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int j = tid + k * kBlock;
        if (j < total) {
            OrderKey key;
            if constexpr (TRACK)
                key = pk[k][6];
            else
                key = (unsigned long long)__double_as_longlong(s_v[j][0]);
            s_key[s_start[slot_rank[k] >> 16] + (slot_rank[k] & 0xffff)] = key;
        }
    }
    __syncthreads();

    // ---- SRD: the cells' draws (axis and the first two Gamma candidates,
    //      which do not depend on the cell's count) on the waves that hold
    //      no cell, while every thread ranks its particles
    static_assert(AT || 2 * NC <= kBlock, "cells and draws on disjoint threads");
    // per cell: the draws (axis, x0, x1, the words of u0 and u1), then the
    // collision's map (axis, u, sc)
    __shared__ double s_dr[AT ? 1 : NC][7];
    if (!AT && tid >= kBlock - NC && !A.inj_axis) {
        const int c = tid - (kBlock - NC);
        const int cx = tc.ox + c % TX, cy = tc.oy + (c / TX) % TY, cz = tc.oz + c / (TX * TY);
        if (s_cnt[c] + (BEADS ? s_bn[c] : 0) > 0 && cx < g.L[0] && cy < g.L[1] && cz < g.L[2]) {
            const long long gid = cx + (long long)g.L[0] * (cy + (long long)g.L[1] * cz);
            double ax[3], x0, x1;
            uint32_t w0, w1;
            cell_draws_words(A.seed, A.step, (uint64_t)gid, ax, x0, w0, x1, w1);
            s_dr[c][0] = ax[0];
            s_dr[c][1] = ax[1];
            s_dr[c][2] = ax[2];
            s_dr[c][3] = x0;
            s_dr[c][4] = x1;
            s_dr[c][5] = __longlong_as_double((long long)(((unsigned long long)w0 << 32) | w1));
        }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int j = tid + k * kBlock;
        if (j < total) {
            const int c = slot_rank[k] >> 16, a = slot_rank[k] & 0xffff;
            const OrderKey *kl = s_key + s_start[c];
            const int n = s_cnt[c];
            const OrderKey kj = kl[a];
            int r = 0;
#pragma unroll 4
            for (int i = 0; i < n; ++i) {
                const OrderKey kq = kl[i];
                r += (kq < kj || (kq == kj && i < a)) ? 1 : 0;
            }
            s_sorted[s_start[c] + r] = (unsigned short)j;
            if constexpr (AT)
                s_rank[j] = (unsigned short)r;
        }
    }
    __syncthreads();

    if constexpr (!AT) {
    // ---- per cell: the fluid's sums in the stable order (mpcd.cpp:158-172,
    //      unit masses), each of Σvx, Σvy, Σvz and Σ|v|^2 a chain of its
    //      own on a thread of its own, in the same order and operations as
    //      one thread adding all four
    double *s_fsum = reinterpret_cast<double *>(s_key);  // [4][NC]
    for (int w = tid; w < 4 * NC; w += kBlock) {
        const int a = w / NC, c = w % NC;
        const unsigned short *lst = s_sorted + s_start[c];
        const int nfl = s_cnt[c];
        double sum = 0.0;
        if (a < 3) {
            for (int i = 0; i < nfl; ++i)
                sum += s_v[lst[i]][a];
        } else {
            for (int i = 0; i < nfl; ++i) {
                const double *v = s_v[lst[i]];
                sum += v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
            }
        }
        s_fsum[w] = sum;
    }
    __syncthreads();
    // ---- per cell: the beads' and virtual particles' sums after the fluid's,
    //      thermostat, rotation, new velocities (mpcd.cpp:158-236)
    if (tid < NC) {
        const int c = tid;
        const int lx = c % TX, ly = (c / TX) % TY, lz = c / (TX * TY);
        const int cx = tc.ox + lx, cy = tc.oy + ly, cz = tc.oz + lz;
        const int nfl = s_cnt[c];
        const int nbd = BEADS ? s_bn[c] : 0;
        int num = nfl + nbd;
        if (num > 0 && cx < g.L[0] && cy < g.L[1] && cz < g.L[2]) {
            const long long gid = cx + (long long)g.L[0] * (cy + (long long)g.L[1] * cz);
            const unsigned short *lst = s_sorted + s_start[c];
            double ux = s_fsum[c], uy = s_fsum[NC + c], uz = s_fsum[2 * NC + c], le = s_fsum[3 * NC + c];
            int masse = nfl;
            if constexpr (BEADS) {
                for (int i = s_bfirst[c]; i < s_bfirst[c] + nbd; ++i) {
                    const double *v = A.b_vel4 + 4 * (long long)i;
                    const double m = v[3];
                    ux += v[0] * m;
                    uy += v[1] * m;
                    uz += v[2] * m;
                    masse += (int)m;
                    le += m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                }
            }
            if (A.walls) {
                // bottom and top layer cells below rho: rho - num virtual
                // particles of momentum sqrt((rho - num) kT) N(0,1) per axis
                // (mpcd.cpp:175-198; the cell test as written there)
                const long long layer = (long long)g.L[0] * g.L[1];
                if ((gid < layer || gid > layer * (g.L[2] - 1)) && num < A.rho) {
                    const int mass_diff = A.rho - num;
                    double nr[3];
                    if (A.inj_virt) {
                        nr[0] = A.inj_virt[3 * gid];
                        nr[1] = A.inj_virt[3 * gid + 1];
                        nr[2] = A.inj_virt[3 * gid + 2];
                    } else {
                        virt_normals(A.seed, A.step, (uint64_t)gid, nr);
                    }
                    const double sv = sqrt(mass_diff * A.kT);
                    const double v0 = sv * nr[0], v1 = sv * nr[1], v2 = sv * nr[2];
                    ux += v0;
                    uy += v1;
                    uz += v2;
                    masse += mass_diff;
                    num += mass_diff;
                    le += (v0 * v0 + v1 * v1 + v2 * v2) / mass_diff;
                }
            }
            // rigid sphere: of rho candidate points in a cell it cuts, those
            // inside are virtual particles moving with it, plus noise
            double dcen[3] = {0.0, 0.0, 0.0};
            bool cut = false;
            if constexpr (SPHERE) {
                const int cc[3] = {cx, cy, cz};
                double q2 = 0.0;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    dcen[k] = min_image((cc[k] + 0.5 - A.a[k]) - A.sphere->X[k], g.L[k]);
                    const double q = fabs(dcen[k]) - 0.5;
                    if (q > 0.0)
                        q2 += q * q;
                }
                cut = q2 < A.sphere->R * A.sphere->R;
                if (cut) {
                    for (int q = 0; q < A.rho; ++q) {
                        const SphereVirtual sv = sphere_virtual(A, gid, dcen[0], dcen[1], dcen[2], q);
                        if (!sv.in)
                            continue;
                        ux += sv.v0;
                        uy += sv.v1;
                        uz += sv.v2;
                        masse += 1;
                        num += 1;
                        le += sv.v0 * sv.v0 + sv.v1 * sv.v1 + sv.v2 * sv.v2;
                    }
                }
            }
            ux /= masse;
            uy /= masse;
            uz /= masse;
            const double rel = le - (ux * ux + uy * uy + uz * uz) * masse;
            double R[9], ax[3], G = 0.0;
            const bool thermo = A.thermostat && rel > 0 && num > 1;
            if (A.inj_axis) {
                ax[0] = A.inj_axis[3 * gid];
                ax[1] = A.inj_axis[3 * gid + 1];
                ax[2] = A.inj_axis[3 * gid + 2];
                if (thermo)
                    G = A.inj_gamma[gid];
            } else {
                ax[0] = s_dr[c][0];
                ax[1] = s_dr[c][1];
                ax[2] = s_dr[c][2];
                if (thermo) {
                    const unsigned long long uw = (unsigned long long)__double_as_longlong(s_dr[c][5]);
                    G = philox_gamma_from2(A.seed, A.step, (uint64_t)gid, num, s_dr[c][3], u32_open01((uint32_t)(uw >> 32)),
                                           s_dr[c][4], u32_open01((uint32_t)uw));
                }
            }
            const double sc = scale_rotation(A, G, ax, num, rel, D, R);
            // the map every particle of the cell applies in the stream pass:
            // v' = u + sc R (v - u), R from the axis (rotmat)
            s_dr[c][0] = ax[0];
            s_dr[c][1] = ax[1];
            s_dr[c][2] = ax[2];
            s_dr[c][3] = ux;
            s_dr[c][4] = uy;
            s_dr[c][5] = uz;
            s_dr[c][6] = sc;
            if constexpr (BEADS) {
                for (int i = s_bfirst[c]; i < s_bfirst[c] + nbd; ++i) {
                    const double *v = A.b_vel4 + 4 * (long long)i;
                    const double m = v[3];
                    double n0, n1, n2;
                    srd_rotate(R, ux, uy, uz, sc, v[0], v[1], v[2], n0, n1, n2);
                    double *o = A.b_out4 + 4 * (long long)A.b_idx[i];
                    o[0] = n0;
                    o[1] = n1;
                    o[2] = n2;
                    o[3] = m;
                    le -= m * (n0 * n0 + n1 * n1 + n2 * n2);
                }
            }
            if (cut) {  // the virtual particles' momentum change goes to the sphere
                double J[3] = {0.0, 0.0, 0.0}, T[3] = {0.0, 0.0, 0.0};
                for (int q = 0; q < A.rho; ++q) {
                    const SphereVirtual sv = sphere_virtual(A, gid, dcen[0], dcen[1], dcen[2], q);
                    if (!sv.in)
                        continue;
                    const double d[3] = {sv.d0, sv.d1, sv.d2}, v[3] = {sv.v0, sv.v1, sv.v2};
                    double n0, n1, n2;
                    srd_rotate(R, ux, uy, uz, sc, v[0], v[1], v[2], n0, n1, n2);
                    const double Jq[3] = {n0 - v[0], n1 - v[1], n2 - v[2]};
#pragma unroll
                    for (int k = 0; k < 3; ++k)
                        J[k] += Jq[k];
                    T[0] += d[1] * Jq[2] - d[2] * Jq[1];
                    T[1] += d[2] * Jq[0] - d[0] * Jq[2];
                    T[2] += d[0] * Jq[1] - d[1] * Jq[0];
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    atomicAdd(&s_sph[k], J[k]);
                    atomicAdd(&s_sph[3 + k], T[k]);
                }
            }
            atomicAdd(&s_dred[0], le);  // the fluid's energy after comes off in the stream pass
            if constexpr (TRACK) {
                if (A.dbg_cell) {
                    for (int i = 0; i < nfl; ++i) {
                        const int id = s_id[lst[i]];
                        A.dbg_cell[id] = (int)gid;
                        A.dbg_rank[id] = i;
                    }
                }
            }
        }
    }

    } else {
        // ---- MPC-AT(+a), absent in the reference (Noguchi, Kikuchi, Gompper,
        //      EPL 78 (2007) 10005): v' = u + xi_i - (sum m xi)/M
        //      [+ omega x r_i], xi_i = sqrt(kT/m_i) N(0,1)^3.  Each particle's
        //      rank in its cell's stable order is set above.
        // per particle: random velocity (Philox of (cell, step, rank), or
        // injected by id) and cell-local coordinate
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int j = tid + k * kBlock;
            if (j < total) {
                const int lc = slot_rank[k] >> 16;
                const int cx = tc.ox + lc % TX, cy = tc.oy + (lc / TX) % TY, cz = tc.oz + lc / (TX * TY);
                const long long gid = cx + (long long)g.L[0] * (cy + (long long)g.L[1] * cz);
                double nr[3];
                if constexpr (TRACK) {
                    if (A.inj_ran) {
                        const long long id = (long long)pk[k][6];
                        nr[0] = A.inj_ran[3 * id];
                        nr[1] = A.inj_ran[3 * id + 1];
                        nr[2] = A.inj_ran[3 * id + 2];
                    } else {
                        ran_normals(A.seed, A.step, (uint64_t)gid, s_rank[j], nr);
                    }
                } else {
                    ran_normals(A.seed, A.step, (uint64_t)gid, s_rank[j], nr);
                }
                const double sk = sqrt(A.kT / 1.0);
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    s_xi[j][a] = sk * nr[a];
                    s_rl[j][a] = (float)local_coord(djoin(pk[k][2 * a], pk[k][2 * a + 1]), A.a[a]);
                }
            }
        }
        __syncthreads();
        // per cell, serially in the stable order (fluid, then beads)
        if (tid < NC) {
            const int c = tid;
            const int cx = tc.ox + c % TX, cy = tc.oy + (c / TX) % TY, cz = tc.oz + c / (TX * TY);
            const int nfl = s_cnt[c];
            const int nbd = BEADS ? s_bn[c] : 0;
            if (nfl + nbd > 0 && cx < g.L[0] && cy < g.L[1] && cz < g.L[2]) {
                const long long gid = cx + (long long)g.L[0] * (cy + (long long)g.L[1] * cz);
                const unsigned short *lst = s_sorted + s_start[c];
                // a bead's random velocity and local coordinate
                auto bead_xi_r = [&](int i, double xi[3], double r[3]) {
                    const double m = A.b_vel4[4 * (long long)i + 3];
                    const int bi = A.b_idx[i];
                    double nr[3];
                    if (A.inj_ran) {
                        const long long at = 3 * (A.nf_total + (long long)bi);
                        nr[0] = A.inj_ran[at];
                        nr[1] = A.inj_ran[at + 1];
                        nr[2] = A.inj_ran[at + 2];
                    } else {
                        ran_normals(A.seed, A.step, (uint64_t)gid, (uint64_t)(nfl + i - s_bfirst[c]), nr);
                    }
                    const double sk = sqrt(A.kT / m);
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        xi[a] = sk * nr[a];
                        r[a] = local_coord(A.b_pos3[3 * (long long)bi + a], A.a[a]);
                    }
                };
                double M = 0.0, P[3] = {0.0, 0.0, 0.0}, X[3] = {0.0, 0.0, 0.0}, C[3] = {0.0, 0.0, 0.0}, le = 0.0;
                auto add1 = [&](const double *v, double m, const double *xi, const double *r) {
                    M += m;
#pragma unroll
                    for (int a = 0; a < 3; ++a) {
                        P[a] += m * v[a];
                        X[a] += m * xi[a];
                        C[a] += m * r[a];
                    }
                    le += m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                };
                for (int i = 0; i < nfl; ++i) {
                    const int j = lst[i];
                    const double r[3] = {(double)s_rl[j][0], (double)s_rl[j][1], (double)s_rl[j][2]};
                    add1(s_v[j], 1.0, s_xi[j], r);
                }
                if constexpr (BEADS) {
                    for (int i = s_bfirst[c]; i < s_bfirst[c] + nbd; ++i) {
                        double xi[3], r[3];
                        bead_xi_r(i, xi, r);
                        add1(A.b_vel4 + 4 * (long long)i, A.b_vel4[4 * (long long)i + 3], xi, r);
                    }
                }
                double u[3], xm[3], w[3] = {0.0, 0.0, 0.0};
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    u[a] = P[a] / M;
                    xm[a] = X[a] / M;
                    C[a] = C[a] / M;
                }
                if constexpr (RULE == kRuleAtA) {
                    double Lc[3] = {0.0, 0.0, 0.0}, I[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
                    auto add2 = [&](const double *v, double m, const double *xi, const double *r0) {
                        const double ri[3] = {r0[0] - C[0], r0[1] - C[1], r0[2] - C[2]};
                        const double d0 = v[0] - xi[0], d1 = v[1] - xi[1], d2 = v[2] - xi[2];
                        Lc[0] += m * (ri[1] * d2 - ri[2] * d1);
                        Lc[1] += m * (ri[2] * d0 - ri[0] * d2);
                        Lc[2] += m * (ri[0] * d1 - ri[1] * d0);
                        I[0] += m * (ri[1] * ri[1] + ri[2] * ri[2]);
                        I[1] += m * (ri[0] * ri[0] + ri[2] * ri[2]);
                        I[2] += m * (ri[0] * ri[0] + ri[1] * ri[1]);
                        I[3] -= m * (ri[0] * ri[1]);
                        I[4] -= m * (ri[0] * ri[2]);
                        I[5] -= m * (ri[1] * ri[2]);
                    };
                    for (int i = 0; i < nfl; ++i) {
                        const int j = lst[i];
                        const double r[3] = {(double)s_rl[j][0], (double)s_rl[j][1], (double)s_rl[j][2]};
                        add2(s_v[j], 1.0, s_xi[j], r);
                    }
                    if constexpr (BEADS) {
                        for (int i = s_bfirst[c]; i < s_bfirst[c] + nbd; ++i) {
                            double xi[3], r[3];
                            bead_xi_r(i, xi, r);
                            add2(A.b_vel4 + 4 * (long long)i, A.b_vel4[4 * (long long)i + 3], xi, r);
                        }
                    }
                    at_omega(I, Lc, w);
                }
                // new velocities
                auto newv = [&](const double *xi, const double *r0, double n[3]) {
                    if constexpr (RULE == kRuleAtA) {
                        const double ri[3] = {r0[0] - C[0], r0[1] - C[1], r0[2] - C[2]};
                        n[0] = ((u[0] + xi[0]) - xm[0]) + (w[1] * ri[2] - w[2] * ri[1]);
                        n[1] = ((u[1] + xi[1]) - xm[1]) + (w[2] * ri[0] - w[0] * ri[2]);
                        n[2] = ((u[2] + xi[2]) - xm[2]) + (w[0] * ri[1] - w[1] * ri[0]);
                    } else {
                        n[0] = (u[0] + xi[0]) - xm[0];
                        n[1] = (u[1] + xi[1]) - xm[1];
                        n[2] = (u[2] + xi[2]) - xm[2];
                    }
                };
                double ea = 0.0, m0 = 0.0, m1 = 0.0, m2 = 0.0;
                for (int i = 0; i < nfl; ++i) {
                    const int j = lst[i];
                    const double r[3] = {(double)s_rl[j][0], (double)s_rl[j][1], (double)s_rl[j][2]};
                    double n[3];
                    newv(s_xi[j], r, n);
                    s_v[j][0] = n[0];
                    s_v[j][1] = n[1];
                    s_v[j][2] = n[2];
                    le -= 1.0 * (n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                    ea += n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
                    m0 += n[0];
                    m1 += n[1];
                    m2 += n[2];
                }
                if constexpr (BEADS) {
                    for (int i = s_bfirst[c]; i < s_bfirst[c] + nbd; ++i) {
                        double xi[3], r[3], n[3];
                        bead_xi_r(i, xi, r);
                        newv(xi, r, n);
                        const double m = A.b_vel4[4 * (long long)i + 3];
                        double *o = A.b_out4 + 4 * (long long)A.b_idx[i];
                        o[0] = n[0];
                        o[1] = n[1];
                        o[2] = n[2];
                        o[3] = m;
                        le -= m * (n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                    }
                }
                atomicAdd(&s_dred[0], le);
                atomicAdd(&s_dred[1], ea);
                atomicAdd(&s_dred[2], m0);
                atomicAdd(&s_dred[3], m1);
                atomicAdd(&s_dred[4], m2);
                if constexpr (TRACK) {
                    if (A.dbg_cell) {
                        for (int i = 0; i < nfl; ++i) {
                            const int id = s_id[lst[i]];
                            A.dbg_cell[id] = (int)gid;
                            A.dbg_rank[id] = i;
                        }
                    }
                }
            }
        }
    }
    __syncthreads();

    // ---- (SRD) each particle's new velocity from its cell's map, in the
    //      operation order of mpcd.cpp:214-231; stream (mpcd.cpp:340-377),
    //      rebin under the next shift, write
    double acc_e = 0.0, acc_p[3] = {0.0, 0.0, 0.0};  // SRD: the fluid's energy and momentum after
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if ((tid & ~63) + k * kBlock >= total)  // no record in this wave: skip its wave-wide steps
            break;
        const int j = tid + k * kBlock;
        int dest = -1;
        uint32_t r[W];
        if (j < total) {
            double n0, n1, n2;
            if constexpr (AT) {
                n0 = s_v[j][0];
                n1 = s_v[j][1];
                n2 = s_v[j][2];
            } else {
                const double *mp = s_dr[slot_rank[k] >> 16];
                const double ax[3] = {mp[0], mp[1], mp[2]};
                double R[9];
                rotmat(ax, A.calpha, A.salpha, R);
                srd_rotate(R, mp[3], mp[4], mp[5], mp[6], s_v[j][0], s_v[j][1], s_v[j][2], n0, n1, n2);
                acc_e += n0 * n0 + n1 * n1 + n2 * n2;
                acc_p[0] += n0;
                acc_p[1] += n1;
                acc_p[2] += n2;
            }
            double x = djoin(pk[k][0], pk[k][1]);
            double y = djoin(pk[k][2], pk[k][3]);
            double z = djoin(pk[k][4], pk[k][5]);
            x += n0 * A.h;
            y += n1 * A.h;
            z += n2 * A.h;
            if (SPHERE && sphere_inside(A, x, y, z)) {
                const SphereHit hh = sphere_hit(A, djoin(pk[k][0], pk[k][1]), djoin(pk[k][2], pk[k][3]),
                                                djoin(pk[k][4], pk[k][5]), x, y, z, n0, n1, n2, s_sph, s_dred);
                x = hh.x;
                y = hh.y;
                z = hh.z;
                n0 = hh.n0;
                n1 = hh.n1;
                n2 = hh.n2;
            }
            if (A.walls) {
                const double o0 = n0, o1 = n1, o2 = n2;
                wall_bounce(A.walls, (double)(g.L[2] - 1), A.vw[0], A.vw[1], x, y, z, n0, n1, n2);
                if (n0 != o0 || n1 != o1 || n2 != o2) {  // diagnostics: the velocity after the bounce
                    atomicAdd(&s_dred[1], (n0 * n0 + n1 * n1 + n2 * n2) - (o0 * o0 + o1 * o1 + o2 * o2));
                    atomicAdd(&s_dred[2], n0 - o0);
                    atomicAdd(&s_dred[3], n1 - o1);
                    atomicAdd(&s_dred[4], n2 - o2);
                }
            }
            dsplit(x, r[0], r[1]);
            dsplit(y, r[2], r[3]);
            dsplit(z, r[4], r[5]);
            dsplit(n0, r[6], r[7]);
            dsplit(n1, r[8], r[9]);
            dsplit(n2, r[10], r[11]);
            if constexpr (TRACK) {
                r[12] = pk[k][6];
                r[13] = 0;
            }
            const int c0 = box_axis_near(x + A.a_next[0], g.L[0], g.invL[0]);
            const int c1 = box_axis_near(y + A.a_next[1], g.L[1], g.invL[1]);
            const int c2 = box_axis_near(z + A.a_next[2], g.L[2], g.invL[2]);
            const int s0 = tile_dir(tc.tx, c0 / TX, g.nt[0]);
            const int s1 = tile_dir(tc.ty, c1 / TY, g.nt[1]);
            const int s2 = tile_dir(tc.tz, c2 / TZ, g.nt[2]);
            const bool far = s0 == 2 || s1 == 2 || s2 == 2;
            dest = far ? 27 : ((s0 | s1 | s2) == 0 ? 0 : 1 + dir26(s0, s1, s2));
        }
        place_record<W>(A, s_dst, s_ocnt, dest, r);
    }
    if constexpr (!AT) {
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            acc_e += __shfl_xor(acc_e, off);
            acc_p[0] += __shfl_xor(acc_p[0], off);
            acc_p[1] += __shfl_xor(acc_p[1], off);
            acc_p[2] += __shfl_xor(acc_p[2], off);
        }
        if ((tid & 63) == 0) {
            atomicAdd(&s_dred[0], -acc_e);
            atomicAdd(&s_dred[1], acc_e);
            atomicAdd(&s_dred[2], acc_p[0]);
            atomicAdd(&s_dred[3], acc_p[1]);
            atomicAdd(&s_dred[4], acc_p[2]);
        }
    }
    __syncthreads();
    write_counts(A, tc.t, s_ocnt);
    if (tid == 0) {
        atomicAdd(&D->diss, s_dred[0]);
        atomicAdd(&D->e2, s_dred[1]);
        atomicAdd(&D->mom[0], s_dred[2]);
        atomicAdd(&D->mom[1], s_dred[3]);
        atomicAdd(&D->mom[2], s_dred[4]);
        atomicAdd((unsigned long long *)&D->n, (unsigned long long)total);
        if constexpr (SPHERE) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                atomicAdd(&A.sphere->dP[k], s_sph[k]);
                atomicAdd(&A.sphere->dL[k], s_sph[3 + k]);
            }
        }
    }
}

// ------------------------------------------------------------ entry helpers
// Main pass, one box: a workgroup per tile.  An ensemble of boxes: one
// launch over the tiles of all of them, box b owning workgroups
// [tile_start[b], tile_start[b + 1]), each box with its own arguments.  The
// bead and large passes: a fixed number of workgroups (per box) stride over
// the box's list; the large pass holds deferred_k particles per thread.  The
// per-cell results do not depend on which pass steps a tile or on the order
// in which its particles are held.
__device__ __forceinline__ int box_of_block(const int *__restrict__ tile_start, int nbox)
{
    int lo = 0, hi = nbox - 1;
    const int blk = (int)blockIdx.x;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tile_start[mid] <= blk)
            lo = mid;
        else
            hi = mid - 1;
    }
    return lo;
}

}  // namespace mcm
