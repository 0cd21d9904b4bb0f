/*
 * CPU oracle for the MPCD step -- TEST INFRASTRUCTURE ONLY (see header).
 * Built with -ffp-contract=off: every product and sum rounds on its own.
 */
#include "mpcd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846 /* src/libcilia/abbrev.cpp:8 */

/* =============================================================== MT19937 */
/* MersenneTwister.h:283-297 initialize(), :300-314 reload(), :122-124 twist */
static void mt_reload(orc_mt *m)
{
    enum { N = 624, M = 397 };
    uint32_t *s = m->s;
    for (int i = 0; i < N; ++i) {
        uint32_t y = (s[i] & 0x80000000u) | (s[(i + 1) % N] & 0x7fffffffu);
        uint32_t v = s[(i + M) % N] ^ (y >> 1) ^ ((s[(i + 1) % N] & 1u) ? 0x9908b0dfu : 0u);
        s[i] = v;
    }
    m->left = N;
    m->next = 0;
}

void orc_mt_seed(orc_mt *m, uint32_t seed)
{
    m->s[0] = seed;
    for (int i = 1; i < 624; ++i)
        m->s[i] = 1812433253u * (m->s[i - 1] ^ (m->s[i - 1] >> 30)) + (uint32_t)i;
    mt_reload(m); /* MersenneTwister.h:219: seed() = initialize + reload */
}

/* MersenneTwister.h:179-193 */
uint32_t orc_mt_u32(orc_mt *m)
{
    if (m->left == 0)
        mt_reload(m);
    --m->left;
    uint32_t y = m->s[m->next++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    return y ^ (y >> 18);
}

double orc_mt_rand(orc_mt *m) { return (double)orc_mt_u32(m) * (1.0 / 4294967295.0); }
double orc_mt_dblexc(orc_mt *m) { return ((double)orc_mt_u32(m) + 0.5) * (1.0 / 4294967296.0); }

/* basicf.h:52-62: Box-Muller, cosine branch only */
double orc_normalrand(orc_mt *m)
{
    double x1 = orc_mt_dblexc(m);
    double x2 = orc_mt_dblexc(m);
    return sqrt(-2.0 * log(x1)) * cos(2.0 * ORC_PI * x2);
}

/* Marsaglia-Tsang Gamma(3(nc-1)/2, 1) as src/libcilia/mpcd.cpp:13-43 */
static double mt_gamma_core(double (*normal)(void *), double (*unif)(void *), void *st, int nc)
{
    const double d = 1.5 * nc - 1.833333333333333333333333333333333333333333;
    const double c = 0.333333333333333333333333333333333333333333 / sqrt(d);
    for (;;) {
        double x, v;
        do {
            x = normal(st);
            v = 1.0 + c * x;
        } while (v <= 0.0);
        v = v * v * v;
        double u = unif(st), x2 = x * x, x4 = x2 * x2;
        if (u < 1.0 - 0.0331 * x4)
            return d * v;
        if (log(u) < 0.5 * x2 + d * (1.0 - v + log(v)))
            return d * v;
    }
}

static double mt_normal_cb(void *st) { return orc_normalrand((orc_mt *)st); }
static double mt_unif_cb(void *st) { return orc_mt_dblexc((orc_mt *)st); }

double orc_rgamma_mt(orc_mt *m, int nc) { return mt_gamma_core(mt_normal_cb, mt_unif_cb, m, nc); }

/* mpcd.cpp:276-284: phi = 2 pi rand(), cos(theta) = 2 rand() - 1 */
void orc_axis_mt(orc_mt *m, double ax[3])
{
    double phi = 2 * ORC_PI * orc_mt_rand(m);
    double theta = 2.0 * orc_mt_rand(m) - 1.0;
    double st = sqrt(1.0 - theta * theta);
    ax[0] = st * cos(phi);
    ax[1] = st * sin(phi);
    ax[2] = theta;
}

/* mpcd.cpp:285-293 (Rodrigues matrix with the given cos/sin constants) */
void orc_rotmat(const double ax[3], double c, double s, double R[9])
{
    const double x = ax[0], y = ax[1], z = ax[2];
    R[0] = x * x + (1.0 - x * x) * c;
    R[1] = x * y * (1.0 - c) - z * s;
    R[2] = x * z * (1.0 - c) + y * s;
    R[3] = x * y * (1.0 - c) + z * s;
    R[4] = y * y + (1.0 - y * y) * c;
    R[5] = y * z * (1.0 - c) - x * s;
    R[6] = x * z * (1.0 - c) - y * s;
    R[7] = y * z * (1.0 - c) + x * s;
    R[8] = z * z + (1.0 - z * z) * c;
}

/* ========================================================= Philox4x64-10 */
static inline void mulhilo64(uint64_t a, uint64_t b, uint64_t *hi, uint64_t *lo)
{
    unsigned __int128 p = (unsigned __int128)a * b;
    *hi = (uint64_t)(p >> 64);
    *lo = (uint64_t)p;
}

void orc_philox4x64(const uint64_t ctr_in[4], const uint64_t key_in[2], uint64_t out[4])
{
    uint64_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint64_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B97F4A7C15ull;
            k1 += 0xBB67AE8584CAA73Bull;
        }
        uint64_t hi0, lo0, hi1, lo1;
        mulhilo64(0xD2E7470EE14C6C93ull, c0, &hi0, &lo0);
        mulhilo64(0xCA5A826395121157ull, c2, &hi1, &lo1);
        uint64_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ========================================================= Philox4x32-10
 * The 32-bit-word member of the same family (Salmon, Moraes, Dror, Shaw,
 * SC'11): multipliers 0xD2511F53, 0xCD9E8D57, Weyl increments 0x9E3779B9,
 * 0xBB67AE85.  The FP32-bulk form's cell draws use it on the device. */
void orc_philox4x32(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4])
{
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n1 = (uint32_t)p1;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1, n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

double orc_u64_to_open01(uint64_t x) { return ((double)(x >> 11) + 0.5) * (1.0 / 9007199254740992.0); }

#define ORC_TAG_CELL 0x43454c4cull
#define ORC_TAG_SHIFT 0x53484654ull
#define ORC_KEY_HI 0x4d504344ull

/* Per-cell uniforms of the production stream: block j of (cell, step) is
 * Philox4x64-10 of counter {cell, step, j, 'CELL'}; its four 64-bit words
 * give eight 32-bit words (low half first), each mapped to (w + 0.5) 2^-32
 * in (0,1) -- the 32-bit resolution of the reference's MT19937 draws
 * (MersenneTwister.h:158-159).  Block 0 holds the rotation axis (U0, U1)
 * and the first two Gamma candidates (U2..U4, U5..U7); candidate k >= 2
 * uses words 3((k-2) mod 2) .. +2 of block 1 + (k-2)/2. */
static void cell_block(uint64_t seed, uint64_t step, uint64_t cell, uint64_t j, double U[8])
{
    uint64_t key[2] = {seed, ORC_KEY_HI};
    uint64_t ctr[4] = {cell, step, j, ORC_TAG_CELL}, o[4];
    orc_philox4x64(ctr, key, o);
    for (int i = 0; i < 4; ++i) {
        U[2 * i] = ((double)(uint32_t)o[i] + 0.5) * (1.0 / 4294967296.0);
        U[2 * i + 1] = ((double)(uint32_t)(o[i] >> 32) + 0.5) * (1.0 / 4294967296.0);
    }
}

/* Candidate k of a cell's thermostat draw: x = Box-Muller cosine branch of
 * (Ua, Ub) as the reference's normalrand (basicf.h:52-62), u = Uc. */
void orc_philox_candidate(uint64_t seed, uint64_t step, uint64_t cell, uint64_t k, double *x, double *u)
{
    double U[8];
    int b;
    if (k < 2) {
        cell_block(seed, step, cell, 0, U);
        b = 2 + 3 * (int)k;
    } else {
        cell_block(seed, step, cell, 1 + (k - 2) / 2, U);
        b = 3 * (int)((k - 2) % 2);
    }
    *x = sqrt(-2.0 * log(U[b])) * cos(2.0 * ORC_PI * U[b + 1]);
    *u = U[b + 2];
}

/* Production Gamma draw: Marsaglia-Tsang over the cell's candidates in
 * order; a candidate with v <= 0 is skipped (the reference's inner redraw
 * loop, mpcd.cpp:34-37). */
double orc_philox_gamma(uint64_t seed, uint64_t step, uint64_t cell, int nc)
{
    const double d = 1.5 * nc - 1.833333333333333333333333333333333333333333;
    const double c = 0.333333333333333333333333333333333333333333 / sqrt(d);
    for (uint64_t k = 0;; ++k) {
        double x, u;
        orc_philox_candidate(seed, step, cell, k, &x, &u);
        double v = 1.0 + c * x;
        if (v <= 0.0)
            continue;
        v = v * v * v;
        const double x2 = x * x, x4 = x2 * x2;
        if (u < 1.0 - 0.0331 * x4)
            return d * v;
        if (log(u) < 0.5 * x2 + d * (1.0 - v + log(v)))
            return d * v;
    }
}

/* Rotation axis uniform on S^2 from U0, U1 of block 0 (mpcd.cpp:276-284). */
void orc_philox_axis(uint64_t seed, uint64_t step, uint64_t cell, double ax[3])
{
    double U[8];
    cell_block(seed, step, cell, 0, U);
    double phi = 2 * ORC_PI * U[0];
    double theta = 2.0 * U[1] - 1.0;
    double st = sqrt(1.0 - theta * theta);
    ax[0] = st * cos(phi);
    ax[1] = st * sin(phi);
    ax[2] = theta;
}

#define ORC_TAG_VIRT 0x56495254ull

/* Virtual wall momentum of a cell: normals k = 0, 1, 2 from words (0,1),
 * (2,3), (4,5) of block 0 of counter {cell, step, 0, 'VIRT'}, each the
 * Box-Muller cosine branch as normalrand (basicf.h:52-62). */
double orc_philox_virt(uint64_t seed, uint64_t step, uint64_t cell, int k)
{
    uint64_t key[2] = {seed, ORC_KEY_HI};
    uint64_t ctr[4] = {cell, step, 0, ORC_TAG_VIRT}, o[4];
    orc_philox4x64(ctr, key, o);
    double U[8];
    for (int i = 0; i < 4; ++i) {
        U[2 * i] = ((double)(uint32_t)o[i] + 0.5) * (1.0 / 4294967296.0);
        U[2 * i + 1] = ((double)(uint32_t)(o[i] >> 32) + 0.5) * (1.0 / 4294967296.0);
    }
    return sqrt(-2.0 * log(U[2 * k])) * cos(2.0 * ORC_PI * U[2 * k + 1]);
}

#define ORC_TAG_MPAT 0x4d504154ull
#define ORC_TAG_SPHV 0x53504856ull
#define ORC_MAX_SPHV 64

/* Candidate k of cell c's virtual sphere particles: offset u_i = U_i (words
 * 0-2) in the cell, normals from Box-Muller on (U3, U4) (cosine and sine
 * branches) and (U5, U6) (cosine), block {cell, step, k, 'SPHV'}. */
void orc_philox_sphv(uint64_t seed, uint64_t step, uint64_t cell, uint64_t k, double u[3], double nr[3])
{
    uint64_t key[2] = {seed, ORC_KEY_HI};
    uint64_t ctr[4] = {cell, step, k, ORC_TAG_SPHV}, o[4];
    orc_philox4x64(ctr, key, o);
    double U[8];
    for (int i = 0; i < 4; ++i) {
        U[2 * i] = ((double)(uint32_t)o[i] + 0.5) * (1.0 / 4294967296.0);
        U[2 * i + 1] = ((double)(uint32_t)(o[i] >> 32) + 0.5) * (1.0 / 4294967296.0);
    }
    u[0] = U[0];
    u[1] = U[1];
    u[2] = U[2];
    const double r1 = sqrt(-2.0 * log(U[3])), r2 = sqrt(-2.0 * log(U[5]));
    nr[0] = r1 * cos(2.0 * ORC_PI * U[4]);
    nr[1] = r1 * sin(2.0 * ORC_PI * U[4]);
    nr[2] = r2 * cos(2.0 * ORC_PI * U[6]);
}

static inline double min_image(double d, int L) { return d - L * rint(d / L); }

/* Relative position, velocity of the sphere surface/interior point d. */
static inline void sphere_point_velocity(const orc_sphere *sp, const double d[3], double v[3])
{
    v[0] = sp->V[0] + (sp->W[1] * d[2] - sp->W[2] * d[1]);
    v[1] = sp->V[1] + (sp->W[2] * d[0] - sp->W[0] * d[2]);
    v[2] = sp->V[2] + (sp->W[0] * d[1] - sp->W[1] * d[0]);
}

void orc_stream_sphere(double *pos3, double *vel4, long n, double ts, const int L[3], orc_sphere *sp)
{
    const double R2 = sp->R * sp->R;
    double X1[3];
    for (int k = 0; k < 3; ++k)
        X1[k] = sp->X[k] + sp->V[k] * ts;
    for (long i = 0; i < n; ++i) {
        double *x = pos3 + 3 * i, *v = vel4 + 4 * i;
        double x1[3], d1[3];
        for (int k = 0; k < 3; ++k) {
            x1[k] = x[k] + v[k] * ts;
            d1[k] = min_image(x1[k] - X1[k], L[k]);
        }
        if (d1[0] * d1[0] + d1[1] * d1[1] + d1[2] * d1[2] >= R2) {
            x[0] = x1[0];
            x[1] = x1[1];
            x[2] = x1[2];
            continue;
        }
        /* relative motion d(t) = d0 + w t, first crossing of |d| = R */
        double w[3], d0[3];
        for (int k = 0; k < 3; ++k) {
            w[k] = v[k] - sp->V[k];
            d0[k] = d1[k] - w[k] * ts;
        }
        const double qa = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
        const double qb = d0[0] * w[0] + d0[1] * w[1] + d0[2] * w[2];
        const double qc = d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2] - R2;
        double t = 0.0;
        if (qc > 0.0 && qa > 0.0) {
            const double disc = qb * qb - qa * qc;
            t = (-qb - sqrt(disc > 0.0 ? disc : 0.0)) / qa;
            if (t < 0.0) t = 0.0;
            if (t > ts) t = ts;
        }
        double dc[3], us[3], vn[3];
        for (int k = 0; k < 3; ++k)
            dc[k] = d0[k] + w[k] * t;
        sphere_point_velocity(sp, dc, us);
        if (sp->slip) {
            const double r = sqrt(dc[0] * dc[0] + dc[1] * dc[1] + dc[2] * dc[2]);
            const double nh[3] = {dc[0] / r, dc[1] / r, dc[2] / r};
            const double vr = (v[0] - us[0]) * nh[0] + (v[1] - us[1]) * nh[1] + (v[2] - us[2]) * nh[2];
            for (int k = 0; k < 3; ++k)
                vn[k] = v[k] - 2.0 * vr * nh[k];
        } else {
            for (int k = 0; k < 3; ++k)
                vn[k] = 2.0 * us[k] - v[k];
        }
        double J[3];
        for (int k = 0; k < 3; ++k) {
            x[k] = (x[k] + v[k] * t) + vn[k] * (ts - t);
            J[k] = v[k] - vn[k];
            sp->dP[k] += J[k];
            v[k] = vn[k];
        }
        sp->dL[0] += dc[1] * J[2] - dc[2] * J[1];
        sp->dL[1] += dc[2] * J[0] - dc[0] * J[2];
        sp->dL[2] += dc[0] * J[1] - dc[1] * J[0];
    }
}

void orc_sphere_finish(orc_sphere *sp, double ts, const int L[3])
{
    for (int k = 0; k < 3; ++k) {
        sp->X[k] += sp->V[k] * ts;
        sp->X[k] -= L[k] * floor(sp->X[k] / L[k]);
        sp->V[k] += sp->dP[k] / sp->Ms;
        sp->W[k] += sp->dL[k] / sp->Is;
        sp->dP[k] = 0.0;
        sp->dL[k] = 0.0;
    }
}

/* MPC-AT random velocity of the particle of rank r in its cell (FP64 form):
 * Philox4x32-10 of counter {cell, low and high word of the step, r}, key
 * {low word of the seed ^ 'MPAT', high word of the seed}; normals 0 and 1
 * the cosine and sine branches of Box-Muller on words 0 and 1, normal 2
 * the cosine branch on words 2 and 3, uniforms (w + 0.5) 2^-32. */
double orc_philox_ran(uint64_t seed, uint64_t step, uint64_t cell, uint64_t r, int k)
{
    const uint32_t key[2] = {(uint32_t)seed ^ (uint32_t)ORC_TAG_MPAT, (uint32_t)(seed >> 32)};
    const uint32_t ctr[4] = {(uint32_t)cell, (uint32_t)step, (uint32_t)(step >> 32), (uint32_t)r};
    uint32_t o[4];
    orc_philox4x32(ctr, key, o);
    const double u0 = ((double)o[0] + 0.5) * (1.0 / 4294967296.0), u1 = ((double)o[1] + 0.5) * (1.0 / 4294967296.0);
    const double u2 = ((double)o[2] + 0.5) * (1.0 / 4294967296.0), u3 = ((double)o[3] + 0.5) * (1.0 / 4294967296.0);
    if (k == 0)
        return sqrt(-2.0 * log(u0)) * cos(2.0 * ORC_PI * u1);
    if (k == 1)
        return sqrt(-2.0 * log(u0)) * sin(2.0 * ORC_PI * u1);
    return sqrt(-2.0 * log(u2)) * cos(2.0 * ORC_PI * u3);
}

/* Shifts on the 2^-24 grid in (0,1): differences are exact in FP32. */
void orc_philox_shift(uint64_t seed, uint64_t step, double a[3])
{
    uint64_t key[2] = {seed, ORC_KEY_HI};
    uint64_t ctr[4] = {0, step, 0, ORC_TAG_SHIFT}, o[4];
    orc_philox4x64(ctr, key, o);
    for (int k = 0; k < 3; ++k)
        a[k] = (double)(1 + (o[k] >> 40) % 16777215ull) * (1.0 / 16777216.0);
}

/* ============================================================== geometry */
static inline int box_axis(double s, int L)
{
    double f = floor(s);
    int i = (int)(f - floor(s / L) * L);
    /* (s)/L can round up to an integer just below a face when L is not a
     * power of two; the reference would then produce -1 (or L).  Wrap. */
    if (i < 0) i += L;
    if (i >= L) i -= L;
    return i;
}

/* mpcd.cpp:130-138 with anull = 1: ix + Lx*iy + Lx*Ly*iz */
int orc_box(const int L[3], const double a[3], const double x[3])
{
    int ix = box_axis(x[0] + a[0], L[0]);
    int iy = box_axis(x[1] + a[1], L[1]);
    int iz = box_axis(x[2] + a[2], L[2]);
    return ix + L[0] * (iy + L[1] * iz);
}

void orc_box_many(const int L[3], const double a[3], const double *pos3, long n, int *cell)
{
    for (long i = 0; i < n; ++i)
        cell[i] = orc_box(L, a, pos3 + 3 * i);
}

/* mpcd.cpp:340-343 (planar walls and BOUNDS are compiled out) */
/* MPCD::move with SLIP or NOSLIP (mpcd.cpp:340-377): the top wall at
 * z = (Lz - 1) reflects (NOSLIP: about the wall velocity, going back along
 * the path for the time t2 past the wall; SLIP: specular), the bottom wall
 * at z = 0 reverses the whole velocity in both cases. */
void orc_move_walls(double *x, double *v, double ts, int Lz, int walls, double vwx, double vwy)
{
    x[0] += v[0] * ts;
    x[1] += v[1] * ts;
    x[2] += v[2] * ts;
    if (!walls)
        return;
    if (x[2] > (Lz - 1) * 1.0) {
        if (walls == 1) {
            const double t2 = (x[2] - (Lz - 1) * 1.0) / v[2];
            v[0] = 2 * vwx - v[0];
            v[1] = 2 * vwy - v[1];
            v[2] *= (-1);
            x[0] += 2.0 * t2 * (v[0] - vwx);
            x[1] += 2.0 * t2 * (v[1] - vwy);
            x[2] += 2.0 * t2 * v[2];
        } else {
            v[2] *= (-1.0);
            x[2] = -x[2] + (Lz - 1) * 1.0 * 2.0;
        }
    }
    if (x[2] < 0) {
        const double t2 = (x[2] / v[2]);
        v[0] *= (-1.0);
        v[1] *= (-1.0);
        v[2] *= (-1.0);
        x[0] += v[0] * 2.0 * t2;
        x[1] += v[1] * 2.0 * t2;
        x[2] += v[2] * 2.0 * t2;
    }
}

void orc_stream_walls(double *pos3, double *vel4, long n, double ts, int Lz, int walls, double vwx, double vwy)
{
    for (long i = 0; i < n; ++i)
        orc_move_walls(pos3 + 3 * i, vel4 + 4 * i, ts, Lz, walls, vwx, vwy);
}

void orc_stream(double *pos3, const double *vel4, long n, double ts)
{
    for (long i = 0; i < n; ++i) {
        pos3[3 * i + 0] += vel4[4 * i + 0] * ts;
        pos3[3 * i + 1] += vel4[4 * i + 1] * ts;
        pos3[3 * i + 2] += vel4[4 * i + 2] * ts;
    }
}

/* fluid.cpp:55-98 (BOUNDS off): per particle 3 rnd() then 3 normalrand() */
void orc_fluid_create(orc_mt *m, const int L[3], long n, double *pos3, double *vel4)
{
    for (long i = 0; i < n; ++i) {
        pos3[3 * i + 0] = orc_mt_dblexc(m) * L[0] * 1.0;
        pos3[3 * i + 1] = orc_mt_dblexc(m) * L[1] * 1.0;
        pos3[3 * i + 2] = orc_mt_dblexc(m) * L[2] * 1.0;
        vel4[4 * i + 0] = orc_normalrand(m) * sqrt(1.0);
        vel4[4 * i + 1] = orc_normalrand(m) * sqrt(1.0);
        vel4[4 * i + 2] = orc_normalrand(m) * sqrt(1.0);
        vel4[4 * i + 3] = 1.0;
    }
    double cm[3] = {0, 0, 0};
    for (long i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k)
            cm[k] += vel4[4 * i + k];
    for (int k = 0; k < 3; ++k)
        cm[k] /= (double)n;
    for (long i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k)
            vel4[4 * i + k] -= cm[k];
}

/* Fluid::create (fluid.cpp:55-98) on the production's counter-based
 * stream: particle i takes Philox4x64-10 blocks {i, 0, 0, 'FLUD'} (three
 * position uniforms, words 0-2, times the extent of each axis) and
 * {i, 1, 0, 'FLUD'} (Box-Muller: r0 = sqrt(-2 ln U0) with angle 2 pi U1
 * gives vx = r0 cos, vy = r0 sin; r1 from U2 with angle 2 pi U3 gives
 * vz = r1 cos), velocities times sqrt(kT), mass 1; then the mean velocity
 * is removed (summed here in index order). */
#define ORC_TAG_FLUID 0x464c5544ull
void orc_fluid_create_philox(uint64_t seed, const double ext[3], double kT, long n, double *pos3, double *vel4)
{
    const uint64_t key[2] = {seed, ORC_KEY_HI};
    const double sk = sqrt(kT);
    for (long i = 0; i < n; ++i) {
        uint64_t c0[4] = {(uint64_t)i, 0, 0, ORC_TAG_FLUID}, c1[4] = {(uint64_t)i, 1, 0, ORC_TAG_FLUID};
        uint64_t p[4], q[4];
        orc_philox4x64(c0, key, p);
        orc_philox4x64(c1, key, q);
        for (int k = 0; k < 3; ++k)
            pos3[3 * i + k] = orc_u64_to_open01(p[k]) * ext[k];
        const double r0 = sqrt(-2.0 * log(orc_u64_to_open01(q[0])));
        const double r1 = sqrt(-2.0 * log(orc_u64_to_open01(q[2])));
        const double t0 = 2.0 * ORC_PI * orc_u64_to_open01(q[1]);
        const double t1 = 2.0 * ORC_PI * orc_u64_to_open01(q[3]);
        vel4[4 * i + 0] = r0 * cos(t0) * sk;
        vel4[4 * i + 1] = r0 * sin(t0) * sk;
        vel4[4 * i + 2] = r1 * cos(t1) * sk;
        vel4[4 * i + 3] = 1.0;
    }
    double cm[3] = {0, 0, 0};
    for (long i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k)
            cm[k] += vel4[4 * i + k];
    for (int k = 0; k < 3; ++k)
        cm[k] /= (double)n;
    for (long i = 0; i < n; ++i)
        for (int k = 0; k < 3; ++k)
            vel4[4 * i + k] -= cm[k];
}

/* ================================================== binning + collision */
void orc_bin(const int *cell, long n, int ncells, long *start, long *perm, int *rank)
{
    memset(start, 0, sizeof(long) * (size_t)(ncells + 1));
    for (long i = 0; i < n; ++i)
        start[cell[i] + 1]++;
    for (int c = 0; c < ncells; ++c)
        start[c + 1] += start[c];
    long *fill = (long *)malloc(sizeof(long) * (size_t)ncells);
    for (int c = 0; c < ncells; ++c)
        fill[c] = start[c];
    for (long i = 0; i < n; ++i) {
        long p = fill[cell[i]]++;
        perm[p] = i;
        if (rank)
            rank[i] = (int)(p - start[cell[i]]);
    }
    free(fill);
}

/* Draws of one occupied cell in the reference order: gamma first (only if
 * rel > 0 and num > 1), then the axis (mpcd.cpp:204-212). */
static double cell_gamma(orc_collide_cfg *cfg, int cid, int num)
{
    switch (cfg->rng) {
    case ORC_RNG_MT: return orc_rgamma_mt(cfg->mt, num);
    case ORC_RNG_INJECT: return cfg->inj_gamma[cid];
    default: return orc_philox_gamma(cfg->seed, cfg->step, (uint64_t)cid, num);
    }
}

static void cell_axis(orc_collide_cfg *cfg, int cid, double ax[3])
{
    switch (cfg->rng) {
    case ORC_RNG_MT: orc_axis_mt(cfg->mt, ax); break;
    case ORC_RNG_INJECT: memcpy(ax, cfg->inj_axis + 3 * (long)cid, 3 * sizeof(double)); break;
    default: orc_philox_axis(cfg->seed, cfg->step, (uint64_t)cid, ax); break;
    }
}

/* Cell-local coordinate of a position: (x + a) - floor(x + a) per axis, the
 * offset inside its shifted cell, rounded to FP32 (the lever arm of the
 * angular-momentum correction; both forms use this value). */
static void local_coord(const double *x, const double a[3], double s[3])
{
    for (int k = 0; k < 3; ++k) {
        const double t = x[k] + a[k];
        s[k] = (double)(float)(t - floor(t));
    }
}

/* omega = I^-1 L for the cell's inertia tensor about its centre of mass
 * (I = [Ixx Ixy Ixz; Ixy Iyy Iyz; Ixz Iyz Izz], off-diagonals signed):
 * adjugate over determinant; when the particles are (nearly) collinear,
 * I = lambda (1 - e e^T) with lambda = tr(I)/2 and L is normal to e, so
 * omega = L / lambda; zero when the cell has no extent. */
void orc_at_omega(const double I[6], const double L[3], double w[3])
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

/* One cell under MPC-AT(+a), particles p of the cell in insertion order
 * (fluid then beads): v' = u + xi_i - (sum m xi)/M [+ omega x r_i], with
 * xi_i = sqrt(kT/m_i) N(0,1)^3, r_i the local coordinate minus the cell's
 * centre of mass, omega = I^-1 sum m r x (v - xi).  Returns the cell's
 * kinetic energy before minus after. */
static double at_cell(orc_collide_cfg *cfg, int cid, const double a[3], const long *pp, long num,
                      const double *pos3, double *vel4, long nf, const double *bpos3, double *bvel4)
{
    double *xi = (double *)malloc(sizeof(double) * 3 * (size_t)num);
    double *r = (double *)malloc(sizeof(double) * 3 * (size_t)num);
    double M = 0.0, P[3] = {0.0, 0.0, 0.0}, X[3] = {0.0, 0.0, 0.0}, C[3] = {0.0, 0.0, 0.0}, le = 0.0;
    for (long q = 0; q < num; ++q) {
        const long i = pp[q];
        double *v = i < nf ? vel4 + 4 * i : bvel4 + 4 * (i - nf);
        const double m = v[3];
        const double sk = sqrt(cfg->kT / m);
        for (int k = 0; k < 3; ++k) {
            double nr;
            switch (cfg->rng) {
            case ORC_RNG_MT: nr = orc_normalrand(cfg->mt); break;
            case ORC_RNG_INJECT: nr = cfg->inj_ran[3 * i + k]; break;
            default: nr = orc_philox_ran(cfg->seed, cfg->step, (uint64_t)cid, (uint64_t)q, k); break;
            }
            if (cfg->rec_ran) cfg->rec_ran[3 * i + k] = nr;
            xi[3 * q + k] = sk * nr;
        }
        local_coord(i < nf ? pos3 + 3 * i : bpos3 + 3 * (i - nf), a, r + 3 * q);
        M += m;
        for (int k = 0; k < 3; ++k) {
            P[k] += m * v[k];
            X[k] += m * xi[3 * q + k];
            C[k] += m * r[3 * q + k];
        }
        le += m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    double u[3], xm[3], w[3] = {0.0, 0.0, 0.0};
    for (int k = 0; k < 3; ++k) {
        u[k] = P[k] / M;
        xm[k] = X[k] / M;
        C[k] = C[k] / M;
    }
    if (cfg->rule == ORC_AT_A) {
        double L[3] = {0.0, 0.0, 0.0}, I[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
        for (long q = 0; q < num; ++q) {
            const long i = pp[q];
            const double *v = i < nf ? vel4 + 4 * i : bvel4 + 4 * (i - nf);
            const double m = v[3];
            double *ri = r + 3 * q;
            for (int k = 0; k < 3; ++k)
                ri[k] = ri[k] - C[k];
            const double d0 = v[0] - xi[3 * q], d1 = v[1] - xi[3 * q + 1], d2 = v[2] - xi[3 * q + 2];
            L[0] += m * (ri[1] * d2 - ri[2] * d1);
            L[1] += m * (ri[2] * d0 - ri[0] * d2);
            L[2] += m * (ri[0] * d1 - ri[1] * d0);
            I[0] += m * (ri[1] * ri[1] + ri[2] * ri[2]);
            I[1] += m * (ri[0] * ri[0] + ri[2] * ri[2]);
            I[2] += m * (ri[0] * ri[0] + ri[1] * ri[1]);
            I[3] -= m * (ri[0] * ri[1]);
            I[4] -= m * (ri[0] * ri[2]);
            I[5] -= m * (ri[1] * ri[2]);
        }
        orc_at_omega(I, L, w);
    }
    for (long q = 0; q < num; ++q) {
        const long i = pp[q];
        double *v = i < nf ? vel4 + 4 * i : bvel4 + 4 * (i - nf);
        const double *ri = r + 3 * q;
        const double *x = xi + 3 * q;
        if (cfg->rule == ORC_AT_A) {
            v[0] = ((u[0] + x[0]) - xm[0]) + (w[1] * ri[2] - w[2] * ri[1]);
            v[1] = ((u[1] + x[1]) - xm[1]) + (w[2] * ri[0] - w[0] * ri[2]);
            v[2] = ((u[2] + x[2]) - xm[2]) + (w[0] * ri[1] - w[1] * ri[0]);
        } else {
            v[0] = (u[0] + x[0]) - xm[0];
            v[1] = (u[1] + x[1]) - xm[1];
            v[2] = (u[2] + x[2]) - xm[2];
        }
        le -= v[3] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    }
    free(xi);
    free(r);
    return le;
}

long double orc_collide(orc_collide_cfg *cfg, const double a[3],
                        const double *pos3, double *vel4, long nf,
                        const double *bpos3, double *bvel4, long nb)
{
    const int *L = cfg->L;
    const int ncells = L[0] * L[1] * L[2];
    const long n = nf + nb;
    int *cell = (int *)malloc(sizeof(int) * (size_t)(n ? n : 1));
    long *start = (long *)malloc(sizeof(long) * (size_t)(ncells + 1));
    long *perm = (long *)malloc(sizeof(long) * (size_t)(n ? n : 1));
    for (long i = 0; i < nf; ++i)
        cell[i] = orc_box(L, a, pos3 + 3 * i);
    for (long j = 0; j < nb; ++j)
        cell[nf + j] = orc_box(L, a, bpos3 + 3 * j);
    orc_bin(cell, n, ncells, start, perm, NULL);
    cfg->rel_violations = 0;
    long double diss = 0.0L;
    for (int cid = 0; cid < ncells; ++cid) {
        if (cfg->rule != ORC_SRD) {
            if (start[cid + 1] > start[cid])
                diss += at_cell(cfg, cid, a, perm + start[cid], start[cid + 1] - start[cid], pos3, vel4, nf, bpos3,
                                bvel4);
            continue;
        }
        if (cfg->rec_gamma) cfg->rec_gamma[cid] = NAN;
        if (cfg->rec_axis) cfg->rec_axis[3 * cid] = cfg->rec_axis[3 * cid + 1] = cfg->rec_axis[3 * cid + 2] = NAN;
        double ux = 0.0, uy = 0.0, uz = 0.0, le = 0.0;
        int num = 0, masse = 0;
        int nsv = 0;
        double svd[ORC_MAX_SPHV][3], svv[ORC_MAX_SPHV][3];
        for (long p = start[cid]; p < start[cid + 1]; ++p) {
            long i = perm[p];
            double *v = i < nf ? vel4 + 4 * i : bvel4 + 4 * (i - nf);
            const double m = v[3];
            ux += v[0] * m;
            uy += v[1] * m;
            uz += v[2] * m;
            masse += (int)m;
            le += m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        }
        num = (int)(start[cid + 1] - start[cid]);
        if (cfg->rec_virt) cfg->rec_virt[3 * cid] = cfg->rec_virt[3 * cid + 1] = cfg->rec_virt[3 * cid + 2] = NAN;
        if (num > 0) {
            /* top or bottom layer cells below rho get virtual wall particles
             * (mpcd.cpp:175-198; the cell test as written there) */
            if (cfg->walls && ((cid < L[0] * L[1]) || (cid > L[0] * L[1] * (L[2] - 1))) && num < cfg->rho) {
                const int mass_diff = cfg->rho - num;
                double vi[3];
                for (int k = 0; k < 3; ++k) {
                    double nr;
                    switch (cfg->rng) {
                    case ORC_RNG_MT: nr = orc_normalrand(cfg->mt); break;
                    case ORC_RNG_INJECT: nr = cfg->inj_virt[3 * (long)cid + k]; break;
                    default: nr = orc_philox_virt(cfg->seed, cfg->step, (uint64_t)cid, k); break;
                    }
                    if (cfg->rec_virt) cfg->rec_virt[3 * (long)cid + k] = nr;
                    vi[k] = sqrt(mass_diff * cfg->kT) * nr;
                }
                ux += vi[0];
                uy += vi[1];
                uz += vi[2];
                masse += mass_diff;
                num += mass_diff;
                le += (vi[0] * vi[0] + vi[1] * vi[1] + vi[2] * vi[2]) / mass_diff;
            }
            /* rigid sphere: of rho candidate points in a cell it cuts, those
             * inside are virtual particles moving with it, plus noise */
            if (cfg->sphere) {
                orc_sphere *sp = cfg->sphere;
                const int cc[3] = {cid % L[0], (cid / L[0]) % L[1], cid / (L[0] * L[1])};
                double dcen[3], q2 = 0.0;
                for (int k = 0; k < 3; ++k) {
                    dcen[k] = min_image((cc[k] + 0.5 - a[k]) - sp->X[k], L[k]);
                    const double q = fabs(dcen[k]) - 0.5;
                    if (q > 0.0)
                        q2 += q * q;
                }
                if (q2 < sp->R * sp->R) {
                    for (int kk = 0; kk < cfg->rho && nsv < ORC_MAX_SPHV; ++kk) {
                        double u[3], nr[3], d[3], v[3];
                        orc_philox_sphv(cfg->seed, cfg->step, (uint64_t)cid, (uint64_t)kk, u, nr);
                        if (cfg->rng == ORC_RNG_INJECT)
                            for (int k = 0; k < 3; ++k)
                                nr[k] = sp->inj_sph[3 * ((long)cfg->rho * cid + kk) + k];
                        for (int k = 0; k < 3; ++k)
                            d[k] = dcen[k] + (u[k] - 0.5);
                        if (d[0] * d[0] + d[1] * d[1] + d[2] * d[2] >= sp->R * sp->R)
                            continue;
                        sphere_point_velocity(sp, d, v);
                        const double sk = sqrt(cfg->kT);
                        for (int k = 0; k < 3; ++k) {
                            v[k] += sk * nr[k];
                            svd[nsv][k] = d[k];
                            svv[nsv][k] = v[k];
                        }
                        ux += v[0];
                        uy += v[1];
                        uz += v[2];
                        masse += 1;
                        num += 1;
                        le += v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
                        ++nsv;
                    }
                }
            }
            ux /= masse;
            uy /= masse;
            uz /= masse;
            double rel = le - (ux * ux + uy * uy + uz * uz) * masse;
            double sc = 1.0;
            if (cfg->thermostat && rel > 0 && num > 1) {
                double g = cell_gamma(cfg, cid, num);
                if (cfg->rec_gamma) cfg->rec_gamma[cid] = g;
                sc = sqrt(cfg->kT * g / (rel * 0.5));
            } else if (!(rel > 0 && num > 1)) {
                if (!(rel * rel < 1e-8))
                    cfg->rel_violations++;
            }
            double ax[3], R[9];
            cell_axis(cfg, cid, ax);
            if (cfg->rec_axis) memcpy(cfg->rec_axis + 3 * (long)cid, ax, sizeof ax);
            orc_rotmat(ax, cfg->calpha, cfg->salpha, R);
            for (long p = start[cid]; p < start[cid + 1]; ++p) {
                long i = perm[p];
                double *v = i < nf ? vel4 + 4 * i : bvel4 + 4 * (i - nf);
                double r0 = v[0] - ux, r1 = v[1] - uy, r2 = v[2] - uz;
                v[0] = ux + (R[0] * r0 + R[1] * r1 + R[2] * r2) * sc;
                v[1] = uy + (R[3] * r0 + R[4] * r1 + R[5] * r2) * sc;
                v[2] = uz + (R[6] * r0 + R[7] * r1 + R[8] * r2) * sc;
                le -= v[3] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            }
            /* the virtual particles' momentum change goes to the sphere */
            for (int q = 0; q < nsv; ++q) {
                const double *v = svv[q], *d = svd[q];
                double r0 = v[0] - ux, r1 = v[1] - uy, r2 = v[2] - uz, J[3];
                J[0] = (ux + (R[0] * r0 + R[1] * r1 + R[2] * r2) * sc) - v[0];
                J[1] = (uy + (R[3] * r0 + R[4] * r1 + R[5] * r2) * sc) - v[1];
                J[2] = (uz + (R[6] * r0 + R[7] * r1 + R[8] * r2) * sc) - v[2];
                for (int k = 0; k < 3; ++k)
                    cfg->sphere->dP[k] += J[k];
                cfg->sphere->dL[0] += d[1] * J[2] - d[2] * J[1];
                cfg->sphere->dL[1] += d[2] * J[0] - d[0] * J[2];
                cfg->sphere->dL[2] += d[0] * J[1] - d[1] * J[0];
            }
        }
        diss += le;
    }
    free(cell);
    free(start);
    free(perm);
    return diss;
}

/* ====================================================== FP32-bulk form */
#define Q_SCALE 2097152.0        /* 2^21 */
#define Q_INV (1.0 / 2097152.0)  /* 2^-21 */

#define Q_EXACT 16777216.0       /* 2^24: grid values below 8 in magnitude */

/* Velocity grid of the FP32-bulk form: the nearest multiple of 2^-21 while
 * that multiple is below 8 in magnitude (|round(v 2^21)| < 2^24), plain
 * FP32 rounding above (those FP32 values are multiples of 2^-21 as well). */
static inline float quant_v(double v)
{
    const double r = rint(v * Q_SCALE);
    if (fabs(r) < Q_EXACT)
        return (float)(r * Q_INV);
    return (float)v;
}

/* The same grid for an FP32 value (the per-particle update computes in
 * FP32): v 2^21 is exact in FP32, so this equals quant_v((double)v). */
static inline float quant_f(float v)
{
    const float r = rintf(v * 2097152.0f);
    if (fabsf(r) < 16777216.0f)
        return r * (1.0f / 2097152.0f);
    return v;
}

static inline long long q_of(float w) { return (long long)((double)w * Q_SCALE); }

void orc_f32_encode(const int L[3], const double a[3], const double *pos3,
                    const double *vel4, long n, int *cell, float *off3, float *vel3)
{
    for (long i = 0; i < n; ++i) {
        int c[3];
        for (int k = 0; k < 3; ++k) {
            double s = pos3[3 * i + k] + a[k];
            double f = floor(s);
            c[k] = box_axis(s, L[k]);
            float o = (float)(s - f);
            if (o >= 1.0f) {
                o = 0.0f;
                c[k] = c[k] + 1 == L[k] ? 0 : c[k] + 1;
            }
            off3[3 * i + k] = o;
            vel3[3 * i + k] = quant_v(vel4[4 * i + k]);
        }
        cell[i] = c[0] + L[0] * (c[1] + L[1] * c[2]);
    }
}

void orc_f32_decode(const int L[3], const double a[3], const int *cell,
                    const float *off3, const float *vel3, long n, double *pos3, double *vel4)
{
    for (long i = 0; i < n; ++i) {
        int c = cell[i];
        int cx = c % L[0], cy = (c / L[0]) % L[1], cz = c / (L[0] * L[1]);
        pos3[3 * i + 0] = (double)cx - a[0] + (double)off3[3 * i + 0];
        pos3[3 * i + 1] = (double)cy - a[1] + (double)off3[3 * i + 1];
        pos3[3 * i + 2] = (double)cz - a[2] + (double)off3[3 * i + 2];
        for (int k = 0; k < 3; ++k)
            vel4[4 * i + k] = (double)vel3[3 * i + k];
        vel4[4 * i + 3] = 1.0;
    }
}

static inline uint32_t fbits(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    return u;
}

static inline uint64_t key2w(uint32_t hi, uint32_t lo) { return ((uint64_t)hi << 32) | lo; }

/* Residual owner order of the FP32-bulk collision.  Keys compared as
 * unsigned integers, lexicographically: (bits(vx), bits(vy)) of the velocity
 * before the collision, then (bits(wx'), bits(wy')), (bits(wz'), bits(ox)),
 * (bits(oy), bits(oz)) of the record after it (new velocity, offsets not yet
 * streamed), then the particle id when ids are tracked.  Particles equal in
 * all of them leave identical records, so any of them may be chosen. */
typedef struct {
    uint64_t k[5];
} owner_key;

static int owner_key_greater(const owner_key *a, const owner_key *b, int nk)
{
    for (int s = 0; s < nk; ++s)
        if (a->k[s] != b->k[s])
            return a->k[s] > b->k[s];
    return 0;
}


/* MPC-AT(+a) in the FP32-bulk form (the GPU's step_f32_tile with RULE
 * kRuleAt / kRuleAtA, restated): the fluid's random velocities on the
 * 2^-21 velocity grid, its lever arms on the 2^-24 grid of the offsets,
 * every fluid sum of the cell an exact integer, bead sums in FP64 in the
 * beads' order.  Normals: injected (fluid id i at 3i, bead j at 3(nf + j))
 * or Philox4x32-10 of counter {cell, bits(ox), bits(oy), bits(oz)} for a
 * fluid particle, {cell, 0xffffffff, j, 0xffffffff} for bead j, key {low
 * word of the seed ^ 'F3AT', high word of the seed ^ low word of the step}:
 * a Box-Muller pair (cosine and sine) from words 0 and 1, the cosine branch
 * of words 2 and 3, on 24-bit uniforms (w >> 8) 2^-24 + 2^-25.  The device
 * evaluates log, sqrt, cos and sin in FP32 on its transcendental unit, this
 * restatement with the C library in double: the two agree to the FP32
 * unit's error (about 1e-7 relative), so a particle's random velocity on the
 * 2^-21 grid may differ by a grid step or two. */
#define ORC_TAG_F3AT 0x46334154ull
#define R_SCALE 16777216.0
#define R_INV (1.0 / 16777216.0)
#define RQ_INV (1.0 / 35184372088832.0)
#define RR_INV (1.0 / 281474976710656.0)

static double u24_open01(uint32_t w) { return (double)(w >> 8) * (1.0 / 16777216.0) + 1.0 / 33554432.0; }

static void f3at_normals(uint64_t seed, uint64_t step, uint64_t cell, uint32_t c1, uint32_t c2, uint32_t c3,
                         double n[3])
{
    const uint32_t key[2] = {(uint32_t)seed ^ (uint32_t)ORC_TAG_F3AT, (uint32_t)(seed >> 32) ^ (uint32_t)step};
    const uint32_t ctr[4] = {(uint32_t)cell, c1, c2, c3};
    uint32_t o[4];
    orc_philox4x32(ctr, key, o);
    const double r0 = sqrt(-2.0 * log(u24_open01(o[0]))), p0 = 2.0 * ORC_PI * u24_open01(o[1]);
    const double r1 = sqrt(-2.0 * log(u24_open01(o[2])));
    n[0] = r0 * cos(p0);
    n[1] = r0 * sin(p0);
    n[2] = r1 * cos(2.0 * ORC_PI * u24_open01(o[3]));
}

static void f3at_fluid_xi(const orc_collide_cfg *cfg, int cid, const float *off, long id, int xq[3])
{
    double n[3];
    if (cfg->rng == ORC_RNG_INJECT) {
        for (int k = 0; k < 3; ++k)
            n[k] = cfg->inj_ran[3 * id + k];
    } else {
        f3at_normals(cfg->seed, cfg->step, (uint64_t)cid, fbits(off[0]), fbits(off[1]), fbits(off[2]), n);
    }
    const double sk = sqrt(cfg->kT);
    for (int k = 0; k < 3; ++k)
        xq[k] = (int)rint(sk * n[k] * Q_SCALE);
}

static void f3at_bead_xi(const orc_collide_cfg *cfg, int cid, long j, long nf, double m, double xi[3])
{
    double n[3];
    if (cfg->rng == ORC_RNG_INJECT) {
        for (int k = 0; k < 3; ++k)
            n[k] = cfg->inj_ran[3 * (nf + j) + k];
    } else {
        f3at_normals(cfg->seed, cfg->step, (uint64_t)cid, 0xffffffffu, (uint32_t)j, 0xffffffffu, n);
    }
    const double sk = sqrt(cfg->kT / m);
    for (int k = 0; k < 3; ++k)
        xi[k] = sk * n[k];
}

static inline long long f3at_rq(float off) { return (long long)rint((double)off * R_SCALE); }

long double orc_f32_step(orc_collide_cfg *cfg, const double a[3], const double a_next[3],
                         double ts, int *cell, float *off3, float *vel3, long n,
                         const double *bpos3, double *bvel4, long nb, const int *ids)
{
    const int *L = cfg->L;
    const int ncells = L[0] * L[1] * L[2];
    long *start = (long *)malloc(sizeof(long) * (size_t)(ncells + 1));
    long *perm = (long *)malloc(sizeof(long) * (size_t)(n ? n : 1));
    int *bcell = (int *)malloc(sizeof(int) * (size_t)(nb ? nb : 1));
    long *bstart = (long *)malloc(sizeof(long) * (size_t)(ncells + 1));
    long *bperm = (long *)malloc(sizeof(long) * (size_t)(nb ? nb : 1));
    long long *qold = (long long *)malloc(sizeof(long long) * 3 * (size_t)(n ? n : 1));
    float *vs = (float *)malloc(sizeof(float) * 3 * (size_t)(n ? n : 1)); /* velocity the stream uses */
    int *xqs = (int *)malloc(sizeof(int) * 3 * (size_t)(n ? n : 1));    /* MPC-AT random velocities, grid units */
    orc_bin(cell, n, ncells, start, perm, NULL);
    for (long j = 0; j < nb; ++j)
        bcell[j] = orc_box(L, a, bpos3 + 3 * j);
    orc_bin(bcell, nb, ncells, bstart, bperm, NULL);
    long double diss_q = 0.0L; /* fluid part, units of 2^-42 (exact per cell) */
    long double diss_b = 0.0L; /* bead part */
    cfg->rel_violations = 0;
    const int nkeys = ids ? 5 : 4;
    for (int cid = 0; cid < ncells; ++cid) {
        long nfl = start[cid + 1] - start[cid];
        long nbd = bstart[cid + 1] - bstart[cid];
        int num = (int)(nfl + nbd);
        if (num == 0)
            continue;
        /* exact sums: S1 = sum q, S2 = sum |q|^2 kept as a 40-bit low part
         * and the high part (q = v 2^21, an integer for every grid value) */
        long long S[3] = {0, 0, 0};
        uint64_t S2lo = 0, S2hi = 0;
        for (long p = start[cid]; p < start[cid + 1]; ++p) {
            long i = perm[p];
            uint64_t s = 0;
            for (int k = 0; k < 3; ++k) {
                long long q = q_of(vel3[3 * i + k]);
                qold[3 * i + k] = q;
                S[k] += q;
                s += (uint64_t)(q * q);
            }
            S2lo += s & ((1ull << 40) - 1);
            S2hi += s >> 40;
        }
        double bx = 0, by = 0, bz = 0, ble = 0;
        int bm = 0;
        for (long p = bstart[cid]; p < bstart[cid + 1]; ++p) {
            const double *v = bvel4 + 4 * bperm[p];
            const double m = v[3];
            bx += v[0] * m;
            by += v[1] * m;
            bz += v[2] * m;
            bm += (int)m;
            ble += m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        }
        /* centre-of-mass velocity and relative energy (mpcd.cpp:200-203) */
        const int masse = (int)nfl + bm;
        const double u[3] = {((double)S[0] * Q_INV + bx) / masse, ((double)S[1] * Q_INV + by) / masse,
                             ((double)S[2] * Q_INV + bz) / masse};
        const double s2 = (double)S2hi * 1099511627776.0 + (double)S2lo; /* 2^40 */
        const double le = s2 * (Q_INV * Q_INV) + ble;
        const double rel = le - (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) * masse;
        double sc = 1.0;
        if (cfg->rule != ORC_SRD) {
            /* MPC-AT: no scale factor, no rel check */
        } else if (cfg->thermostat && rel > 0 && num > 1) {
            double g = cell_gamma(cfg, cid, num);
            sc = sqrt(cfg->kT * g / (rel * 0.5));
        } else if (!(rel > 0 && num > 1)) {
            if (!(rel * rel < 1e-8))
                cfg->rel_violations++;
        }
        const int at = cfg->rule != ORC_SRD, ata = cfg->rule == ORC_AT_A;
        double bp[3] = {0, 0, 0}, om[3] = {0, 0, 0};
        float bpf[3] = {0, 0, 0}, omf[3] = {0, 0, 0};
        if (at) {
            /* fluid: random velocities in grid units and the exact sums */
            uint64_t SX[3] = {0, 0, 0}, SR[3] = {0, 0, 0}, SD[3] = {0, 0, 0}, SRR[6] = {0, 0, 0, 0, 0, 0};
            for (long p = start[cid]; p < start[cid + 1]; ++p) {
                long i = perm[p];
                int xq[3];
                f3at_fluid_xi(cfg, cid, off3 + 3 * i, ids ? (long)ids[i] : i, xq);
                uint64_t rq[3], dq[3];
                for (int k = 0; k < 3; ++k) {
                    xqs[3 * i + k] = xq[k];
                    SX[k] += (uint64_t)(long long)xq[k];
                    rq[k] = (uint64_t)f3at_rq(off3[3 * i + k]);
                    dq[k] = (uint64_t)(qold[3 * i + k] - xq[k]);
                    SR[k] += rq[k];
                }
                SD[0] += rq[1] * dq[2] - rq[2] * dq[1];
                SD[1] += rq[2] * dq[0] - rq[0] * dq[2];
                SD[2] += rq[0] * dq[1] - rq[1] * dq[0];
                SRR[0] += rq[0] * rq[0];
                SRR[1] += rq[1] * rq[1];
                SRR[2] += rq[2] * rq[2];
                SRR[3] += rq[0] * rq[1];
                SRR[4] += rq[0] * rq[2];
                SRR[5] += rq[1] * rq[2];
            }
            /* beads in their order (FP64) */
            double bmv[3] = {0, 0, 0}, bX[3] = {0, 0, 0}, bC[3] = {0, 0, 0}, bL[3] = {0, 0, 0};
            double bI[6] = {0, 0, 0, 0, 0, 0};
            int bmi = 0;
            for (long p = bstart[cid]; p < bstart[cid + 1]; ++p) {
                const long j = bperm[p];
                const double *v = bvel4 + 4 * j;
                const double m = v[3];
                double xi[3];
                f3at_bead_xi(cfg, cid, j, n, m, xi);
                bmv[0] += v[0] * m;
                bmv[1] += v[1] * m;
                bmv[2] += v[2] * m;
                bmi += (int)m;
                for (int k = 0; k < 3; ++k)
                    bX[k] += m * xi[k];
                if (ata) {
                    double r[3], d[3];
                    local_coord(bpos3 + 3 * j, a, r);
                    for (int k = 0; k < 3; ++k) {
                        d[k] = v[k] - xi[k];
                        bC[k] += m * r[k];
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
            const double Mc = (double)(nfl + bmi);
            double uu[3], xm[3];
            for (int k = 0; k < 3; ++k) {
                const double S1 = (double)S[k] * Q_INV;
                uu[k] = (S1 + bmv[k]) / Mc;
                xm[k] = ((double)(long long)SX[k] * Q_INV + bX[k]) / Mc;
            }
            if (ata) {
                double Cc[3], Dm[3], Lc[3], I[6];
                for (int k = 0; k < 3; ++k) {
                    Cc[k] = ((double)(long long)SR[k] * R_INV + bC[k]) / Mc;
                    Dm[k] = ((double)S[k] * Q_INV + bmv[k]) - ((double)(long long)SX[k] * Q_INV + bX[k]);
                }
                Lc[0] = ((double)(long long)SD[0] * RQ_INV + bL[0]) - (Cc[1] * Dm[2] - Cc[2] * Dm[1]);
                Lc[1] = ((double)(long long)SD[1] * RQ_INV + bL[1]) - (Cc[2] * Dm[0] - Cc[0] * Dm[2]);
                Lc[2] = ((double)(long long)SD[2] * RQ_INV + bL[2]) - (Cc[0] * Dm[1] - Cc[1] * Dm[0]);
                I[0] = ((double)(long long)(SRR[1] + SRR[2]) * RR_INV + bI[0]) - Mc * (Cc[1] * Cc[1] + Cc[2] * Cc[2]);
                I[1] = ((double)(long long)(SRR[0] + SRR[2]) * RR_INV + bI[1]) - Mc * (Cc[0] * Cc[0] + Cc[2] * Cc[2]);
                I[2] = ((double)(long long)(SRR[0] + SRR[1]) * RR_INV + bI[2]) - Mc * (Cc[0] * Cc[0] + Cc[1] * Cc[1]);
                I[3] = (bI[3] - (double)(long long)SRR[3] * RR_INV) + Mc * (Cc[0] * Cc[1]);
                I[4] = (bI[4] - (double)(long long)SRR[4] * RR_INV) + Mc * (Cc[0] * Cc[2]);
                I[5] = (bI[5] - (double)(long long)SRR[5] * RR_INV) + Mc * (Cc[1] * Cc[2]);
                orc_at_omega(I, Lc, om);
                bp[0] = (uu[0] - xm[0]) - (om[1] * Cc[2] - om[2] * Cc[1]);
                bp[1] = (uu[1] - xm[1]) - (om[2] * Cc[0] - om[0] * Cc[2]);
                bp[2] = (uu[2] - xm[2]) - (om[0] * Cc[1] - om[1] * Cc[0]);
            } else {
                for (int k = 0; k < 3; ++k)
                    bp[k] = uu[k] - xm[k];
            }
            for (int k = 0; k < 3; ++k) {
                bpf[k] = (float)bp[k];
                omf[k] = (float)om[k];
            }
        }
        double ax[3], R[9];
        if (!at) {
            cell_axis(cfg, cid, ax);
            orc_rotmat(ax, cfg->calpha, cfg->salpha, R);
        } else {
            memset(R, 0, sizeof R);
        }
        /* the collision as one affine map v' = M v + b, M = sc R,
         * b = u - M u (mpcd.cpp:214-231 regrouped), with fused multiply-adds:
         * FP64 for the beads, rounded to FP32 for the solvent update */
        double M[9], b[3];
        float Mf[9], bf[3];
        for (int k = 0; k < 9; ++k) {
            M[k] = R[k] * sc;
            Mf[k] = (float)M[k];
        }
        for (int k = 0; k < 3; ++k) {
            double t = M[3 * k + 2] * u[2];
            t = fma(M[3 * k + 1], u[1], t);
            t = fma(M[3 * k], u[0], t);
            b[k] = u[k] - t;
            bf[k] = (float)b[k];
        }
        long long D[3] = {0, 0, 0}; /* sum of q - q' */
        long owner = -1;
        owner_key okey;
        memset(&okey, 0, sizeof okey);
        for (long p = start[cid]; p < start[cid + 1]; ++p) {
            long i = perm[p];
            owner_key key;
            key.k[0] = key2w(fbits(vel3[3 * i]), fbits(vel3[3 * i + 1]));
            const float v0 = vel3[3 * i], v1 = vel3[3 * i + 1], v2 = vel3[3 * i + 2];
            float atv[3] = {0, 0, 0};
            if (at) {
                float rf[3];
                for (int k = 0; k < 3; ++k) {
                    atv[k] = bpf[k] + (float)xqs[3 * i + k] * (1.0f / 2097152.0f);
                    rf[k] = (float)f3at_rq(off3[3 * i + k]) * (1.0f / 16777216.0f);
                }
                if (ata)
                    for (int k = 0; k < 3; ++k) {
                        const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
                        atv[k] = fmaf(omf[k1], rf[k2], atv[k]);
                        atv[k] = fmaf(-omf[k2], rf[k1], atv[k]);
                    }
            }
            for (int k = 0; k < 3; ++k) {
                const float vn = at ? atv[k]
                                    : fmaf(Mf[3 * k], v0, fmaf(Mf[3 * k + 1], v1, fmaf(Mf[3 * k + 2], v2, bf[k])));
                const float w = quant_f(vn);
                D[k] += qold[3 * i + k] - q_of(w);
                vel3[3 * i + k] = w;
                vs[3 * i + k] = w;
            }
            key.k[1] = key2w(fbits(vel3[3 * i]), fbits(vel3[3 * i + 1]));
            key.k[2] = key2w(fbits(vel3[3 * i + 2]), fbits(off3[3 * i]));
            key.k[3] = key2w(fbits(off3[3 * i + 1]), fbits(off3[3 * i + 2]));
            key.k[4] = ids ? (uint64_t)(uint32_t)ids[i] : 0;
            if (owner < 0 || owner_key_greater(&key, &okey, nkeys)) {
                owner = i;
                okey = key;
            }
        }
        /* exact momentum: one particle takes the residual (fluid-only cells)
         * into its stored velocity; it streams with the velocity before the
         * residual, as every other particle streams with its own */
        if (nbd == 0 && owner >= 0)
            for (int k = 0; k < 3; ++k)
                vel3[3 * owner + k] = (float)((double)(q_of(vel3[3 * owner + k]) + D[k]) * Q_INV);
        uint64_t S2n = 0;
        for (long p = start[cid]; p < start[cid + 1]; ++p) {
            long i = perm[p];
            for (int k = 0; k < 3; ++k) {
                long long q = q_of(vel3[3 * i + k]);
                S2n += (uint64_t)(q * q);
            }
        }
        diss_q += (long double)s2 - (long double)S2n;
        /* beads: the FP64 map; then the first bead of the cell takes the
         * cell's momentum balance, so fluid (exact integers) + beads
         * conserve momentum to FP64 roundoff */
        double Pb[3] = {0, 0, 0};
        for (long p = bstart[cid]; p < bstart[cid + 1]; ++p) {
            const long j = bperm[p];
            double *v = bvel4 + 4 * j;
            const double v0 = v[0], v1 = v[1], v2 = v[2];
            if (at) {
                double xi[3], r[3], nv[3];
                f3at_bead_xi(cfg, cid, j, n, v[3], xi);
                if (ata) {
                    local_coord(bpos3 + 3 * j, a, r);
                    for (int k = 0; k < 3; ++k) {
                        const int k1 = (k + 1) % 3, k2 = (k + 2) % 3;
                        nv[k] = (bp[k] + xi[k]) + (om[k1] * r[k2] - om[k2] * r[k1]);
                    }
                } else {
                    for (int k = 0; k < 3; ++k)
                        nv[k] = bp[k] + xi[k];
                }
                for (int k = 0; k < 3; ++k) {
                    v[k] = nv[k];
                    Pb[k] += v[k] * v[3];
                }
                continue;
            }
            for (int k = 0; k < 3; ++k) {
                v[k] = fma(M[3 * k], v0, fma(M[3 * k + 1], v1, fma(M[3 * k + 2], v2, b[k])));
                Pb[k] += v[k] * v[3];
            }
        }
        if (nbd) {
            double *v = bvel4 + 4 * bperm[bstart[cid]];
            const double bb[3] = {bx, by, bz};
            for (int k = 0; k < 3; ++k) {
                const double before = (double)S[k] * Q_INV + bb[k];
                const double after = (double)(S[k] - D[k]) * Q_INV + Pb[k];
                v[k] += (before - after) / v[3];
            }
        }
        double blen = 0;
        for (long p = bstart[cid]; p < bstart[cid + 1]; ++p) {
            const double *v = bvel4 + 4 * bperm[p];
            blen += v[3] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        }
        diss_b += (long double)(ble - blen);
    }
    /* stream + rebin with the next shift, per axis in FP32 */
    const float h = (float)ts;
    float da[3];
    for (int k = 0; k < 3; ++k)
        da[k] = (float)(a_next[k] - a[k]);
    for (long i = 0; i < n; ++i) {
        int c = cell[i];
        int ck[3] = {c % L[0], (c / L[0]) % L[1], c / (L[0] * L[1])};
        for (int k = 0; k < 3; ++k) {
            float t1 = vs[3 * i + k] * h;
            float t2 = off3[3 * i + k] + t1;
            float pk = t2 + da[k];
            float fl = floorf(pk);
            int nk = (int)fl;
            float o = pk - fl;
            if (o >= 1.0f) {
                o = 0.0f;
                nk += 1;
            }
            int cn = (ck[k] + nk) % L[k];
            if (cn < 0) cn += L[k];
            ck[k] = cn;
            off3[3 * i + k] = o;
        }
        cell[i] = ck[0] + L[0] * (ck[1] + L[1] * ck[2]);
    }
    free(start);
    free(perm);
    free(bcell);
    free(bstart);
    free(bperm);
    free(qold);
    free(vs);
    free(xqs);
    return diss_q * (long double)(Q_INV * Q_INV) + diss_b;
}

/* ============================================ OpenMP bucket step (baseline) */
/* A bucket is the reference's std::vector<double *> with 16 entries
 * reserved (mpcd.cpp:76-83): the first 16 live in one arena for all buckets
 * (fast to set up at 10^8 buckets), a bucket that outgrows them moves to
 * its own heap array, doubling, as the vector does. */
typedef struct {
    double **p;
    int size, cap;
    int own; /* p is a heap array of this bucket, not the arena's */
} bucket;

struct orc_omp_sim {
    int L[3];
    int ncells;
    long n;
    double *pos, *vel;
    int T;
    bucket *b; /* T * ncells, thread-major as mpcd.h:54 */
    double **arena;
    double *le;
    orc_mt *mt; /* one per thread (basicf.cpp:14) */
    double a[3];
    long double diss;
};

static inline void bucket_push(bucket *b, double *v)
{
    if (b->size == b->cap) {
        const int cap = 2 * b->cap;
        double **q = (double **)malloc(sizeof(double *) * (size_t)cap);
        memcpy(q, b->p, sizeof(double *) * (size_t)b->size);
        if (b->own)
            free(b->p);
        b->p = q;
        b->cap = cap;
        b->own = 1;
    }
    b->p[b->size++] = v;
}

orc_omp_sim *orc_omp_create(const int L[3], long n, double *pos3, double *vel4, int nthreads, uint32_t seed)
{
    orc_omp_sim *s = (orc_omp_sim *)calloc(1, sizeof *s);
    memcpy(s->L, L, sizeof s->L);
    s->ncells = L[0] * L[1] * L[2];
    s->n = n;
    s->pos = pos3;
    s->vel = vel4;
#ifdef _OPENMP
    s->T = nthreads > 0 ? nthreads : omp_get_max_threads();
#else
    s->T = 1;
#endif
    const long nbk = (long)s->T * s->ncells;
    s->b = (bucket *)calloc((size_t)nbk, sizeof(bucket));
    s->arena = (double **)malloc(sizeof(double *) * 16 * (size_t)nbk);
    /* mpcd.cpp:76-83: every bucket reserves 16 entries up front */
    for (long i = 0; i < nbk; ++i) {
        s->b[i].cap = 16;
        s->b[i].p = s->arena + 16 * i;
    }
    s->le = (double *)calloc((size_t)s->ncells, sizeof(double));
    s->mt = (orc_mt *)malloc(sizeof(orc_mt) * (size_t)s->T);
    for (int t = 0; t < s->T; ++t)
        orc_mt_seed(&s->mt[t], seed + (uint32_t)t);
    for (int k = 0; k < 3; ++k)
        s->a[k] = orc_mt_dblexc(&s->mt[0]);
    for (long i = 0; i < n; ++i)
        bucket_push(&s->b[orc_box(L, s->a, pos3 + 3 * i)], vel4 + 4 * i);
    return s;
}

int orc_omp_threads(orc_omp_sim *s) { return s->T; }

void orc_omp_steps(orc_omp_sim *s, int steps, double tstep)
{
    const int T = s->T, ncells = s->ncells;
    const double calpha = -0.6427876, salpha = 0.76604444;
    for (int it = 0; it < steps; ++it) {
        long double diss = 0.0L;
#pragma omp parallel num_threads(T)
        {
#ifdef _OPENMP
            const int tid = omp_get_thread_num();
#else
            const int tid = 0;
#endif
            orc_mt *mt = &s->mt[tid];
            /* mpcd.cpp:140-242 */
#pragma omp for schedule(static) nowait
            for (int cid = 0; cid < ncells; ++cid) {
                double ux = 0.0, uy = 0.0, uz = 0.0, le = 0.0;
                int num = 0, masse = 0;
                for (int t = 0; t < T; ++t) {
                    bucket *b = &s->b[(long)t * ncells + cid];
                    for (int k = 0; k < b->size; ++k) {
                        double *v = b->p[k];
                        const double m = v[3];
                        ux += v[0] * m;
                        uy += v[1] * m;
                        uz += v[2] * m;
                        masse += (int)m;
                        le += m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                    }
                    num += b->size;
                }
                if (num > 0) {
                    ux /= masse;
                    uy /= masse;
                    uz /= masse;
                    double rel = le - (ux * ux + uy * uy + uz * uz) * masse;
                    double sc = 1.0;
                    if (rel > 0 && num > 1)
                        sc = sqrt(orc_rgamma_mt(mt, num) / (rel * 0.5));
                    double ax[3], R[9];
                    orc_axis_mt(mt, ax);
                    orc_rotmat(ax, calpha, salpha, R);
                    for (int t = 0; t < T; ++t) {
                        bucket *b = &s->b[(long)t * ncells + cid];
                        for (int k = 0; k < b->size; ++k) {
                            double *v = b->p[k];
                            double r0 = v[0] - ux, r1 = v[1] - uy, r2 = v[2] - uz;
                            v[0] = ux + (R[0] * r0 + R[1] * r1 + R[2] * r2) * sc;
                            v[1] = uy + (R[3] * r0 + R[4] * r1 + R[5] * r2) * sc;
                            v[2] = uz + (R[6] * r0 + R[7] * r1 + R[8] * r2) * sc;
                            le -= v[3] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
                        }
                    }
                }
                s->le[cid] = le;
                for (int t = 0; t < T; ++t)
                    s->b[(long)t * ncells + cid].size = 0;
            }
            /* mpcd.cpp:243-246; a double partial per thread, then combined */
            double part = 0.0;
#pragma omp for schedule(static)
            for (int cid = 0; cid < ncells; ++cid)
                part += s->le[cid];
#pragma omp critical
            diss += part;
        }
        s->diss += diss;
        /* mpcd.cpp:249-253 (serial) */
        for (int k = 0; k < 3; ++k)
            s->a[k] = orc_mt_dblexc(&s->mt[0]);
        /* fluid.cpp:152-167 */
#pragma omp parallel num_threads(T)
        {
#ifdef _OPENMP
            const int tid = omp_get_thread_num();
#else
            const int tid = 0;
#endif
#pragma omp for schedule(static) nowait
            for (long i = 0; i < s->n; ++i) {
                double *x = s->pos + 3 * i, *v = s->vel + 4 * i;
                x[0] += v[0] * tstep;
                x[1] += v[1] * tstep;
                x[2] += v[2] * tstep;
                bucket_push(&s->b[(long)tid * ncells + orc_box(s->L, s->a, x)], v);
            }
        }
    }
}

void orc_omp_destroy(orc_omp_sim *s)
{
    if (!s)
        return;
    for (long i = 0; i < (long)s->T * s->ncells; ++i)
        if (s->b[i].own)
            free(s->b[i].p);
    free(s->arena);
    free(s->b);
    free(s->le);
    free(s->mt);
    free(s);
}
