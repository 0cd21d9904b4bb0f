// Device-side building blocks of the MPCD step for gfx950 (CDNA4).
//
// All FP64 arithmetic here is written so that, compiled with
// -ffp-contract=off, every product and sum rounds separately; the CPU oracle
// is built the same way, which is what makes the FP64 form bit-comparable.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcm {

constexpr double kPi = 3.14159265358979323846;  // src/libcilia/abbrev.cpp:8

// ------------------------------------------------------------ Philox4x64-10
// Counter-based generator (Salmon, Moraes, Dror, Shaw, SC'11).  10 rounds,
// multipliers and Weyl increments of the published Philox4x64 definition.
struct U64x4 {
    uint64_t v[4];
};

__device__ __forceinline__ U64x4 philox4x64(uint64_t c0, uint64_t c1, uint64_t c2, uint64_t c3,
                                            uint64_t k0, uint64_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B97F4A7C15ull;
            k1 += 0xBB67AE8584CAA73Bull;
        }
        const uint64_t lo0 = 0xD2E7470EE14C6C93ull * c0;
        const uint64_t hi0 = __umul64hi(0xD2E7470EE14C6C93ull, c0);
        const uint64_t lo1 = 0xCA5A826395121157ull * c2;
        const uint64_t hi1 = __umul64hi(0xCA5A826395121157ull, c2);
        const uint64_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0;
        c1 = lo1;
        c2 = n2;
        c3 = lo0;
    }
    return U64x4{{c0, c1, c2, c3}};
}

// ------------------------------------------------------------ Philox4x32-10
// The same generator on 32-bit words (multipliers and Weyl increments of
// the published Philox4x32 definition): its multiplies are native 32 x 32
// -> 64-bit products, where Philox4x64's are emulated.
struct U32x4 {
    uint32_t v[4];
};

__device__ __forceinline__ U32x4 philox4x32(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3, uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) {
            k0 += 0x9E3779B9u;
            k1 += 0xBB67AE85u;
        }
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
    }
    return U32x4{{c0, c1, c2, c3}};
}

__device__ __forceinline__ double u64_open01(uint64_t x)
{
    return ((double)(x >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

constexpr uint64_t kTagCell = 0x43454c4cull;
constexpr uint64_t kKeyHi = 0x4d504344ull;

// (w + 0.5) 2^-32 in (0,1): the 32-bit resolution of the reference's
// MT19937 draws (MersenneTwister.h:158-159).
__device__ __forceinline__ double u32_open01(uint32_t w) { return ((double)w + 0.5) * (1.0 / 4294967296.0); }

// Block j of a cell's stream at a step: Philox4x64-10 of counter
// {cell, step, j, 'CELL'}, its four words split into eight 32-bit words
// (low half first).  Block 0 holds the rotation axis (words 0, 1) and the
// first two Gamma candidates (words 2-4, 5-7); candidate k >= 2 uses words
// 3((k-2) mod 2) .. +2 of block 1 + (k-2)/2.
__device__ __forceinline__ void cell_block(uint64_t seed, uint64_t step, uint64_t cell, uint64_t j, uint32_t (&w)[8])
{
    const U64x4 o = philox4x64(cell, step, j, kTagCell, seed, kKeyHi);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        w[2 * i] = (uint32_t)o.v[i];
        w[2 * i + 1] = (uint32_t)(o.v[i] >> 32);
    }
}

// Box-Muller cosine branch of (Ua, Ub), the reference's normalrand
// (src/libcilia/basicf.h:52-62); cos(2 pi U) as cospi(2 U).
__device__ __forceinline__ double box_muller(uint32_t a, uint32_t b)
{
    return sqrt(-2.0 * log(u32_open01(a))) * cospi(2.0 * u32_open01(b));
}

// Acceptance test of a candidate for Gamma(3(nc-1)/2): mpcd.cpp:13-43.
__device__ __forceinline__ bool gamma_accept(double d, double c, double x, double u, double &g)
{
    double v = 1.0 + c * x;
    if (v <= 0.0)
        return false;
    v = v * v * v;
    const double x2 = x * x, x4 = x2 * x2;
    if (u < 1.0 - 0.0331 * x4 || log(u) < 0.5 * x2 + d * (1.0 - v + log(v))) {
        g = d * v;
        return true;
    }
    return false;
}

struct GammaCand {
    double x, u;
};

// Candidate k >= 2 of a cell: words 3((k-2) mod 2) .. +2 of block
// 1 + (k-2)/2.  Out of line, one candidate per call, so the rare tail below
// needs few registers.
__device__ __noinline__ GammaCand gamma_cand_tail(uint64_t seed, uint64_t step, uint64_t cell, uint64_t k)
{
    uint32_t w[8];
    cell_block(seed, step, cell, 1 + (k - 2) / 2, w);
    const bool hi = ((k - 2) & 1) != 0;
    const uint32_t a = hi ? w[3] : w[0], b = hi ? w[4] : w[1], c = hi ? w[5] : w[2];
    return GammaCand{box_muller(a, b), u32_open01(c)};
}

// Candidates 2, 3, ... until one is accepted.  Kept out of line: the
// kernels draw the first two candidates with the axis and land here only
// when both are rejected.
__device__ __noinline__ double gamma_tail(uint64_t seed, uint64_t step, uint64_t cell, double d, double c)
{
#pragma unroll 1
    for (uint64_t k = 2;; ++k) {
        const GammaCand cd = gamma_cand_tail(seed, step, cell, k);
        double g;
        if (gamma_accept(d, c, cd.x, cd.u, g))
            return g;
    }
}

// Gamma(3(nc-1)/2, 1) variate of a cell given its first two candidates.
__device__ __forceinline__ double philox_gamma_from2(uint64_t seed, uint64_t step, uint64_t cell, int nc, double x0,
                                                     double u0, double x1, double u1)
{
    const double d = 1.5 * nc - 1.833333333333333333333333333333333333333333;
    const double c = 0.333333333333333333333333333333333333333333 / sqrt(d);
    double g;
    if (gamma_accept(d, c, x0, u0, g) || gamma_accept(d, c, x1, u1, g))
        return g;
    return gamma_tail(seed, step, cell, d, c);
}

// Block 0 of a cell: the rotation axis, uniform on S^2 (phi = 2 pi U0,
// cos(theta) = 2 U1 - 1, src/libcilia/mpcd.cpp:276-284), and the first two
// Gamma candidates (x, u), u as its 32-bit word (u = u32_open01(word)).
__device__ __forceinline__ void cell_draws_words(uint64_t seed, uint64_t step, uint64_t cell, double ax[3],
                                                 double &x0, uint32_t &uw0, double &x1, uint32_t &uw1)
{
    uint32_t w[8];
    cell_block(seed, step, cell, 0, w);
    double sp, cp;
    sincospi(2.0 * u32_open01(w[0]), &sp, &cp);
    const double ct = 2.0 * u32_open01(w[1]) - 1.0;
    const double st = sqrt(1.0 - ct * ct);
    ax[0] = st * cp;
    ax[1] = st * sp;
    ax[2] = ct;
    x0 = box_muller(w[2], w[3]);
    uw0 = w[4];
    x1 = box_muller(w[5], w[6]);
    uw1 = w[7];
}

__device__ __forceinline__ void cell_draws(uint64_t seed, uint64_t step, uint64_t cell, double ax[3], double &x0,
                                           double &u0, double &x1, double &u1)
{
    uint32_t uw0, uw1;
    cell_draws_words(seed, step, cell, ax, x0, uw0, x1, uw1);
    u0 = u32_open01(uw0);
    u1 = u32_open01(uw1);
}

// The cell draws of the FP32-bulk form, whose map is FP32 anyway: two
// Philox4x32-10 blocks of counter {cell, step low, step high, 'F3CL' + j}
// and key {seed low, seed high}, with FP32 arithmetic on the hardware
// transcendental unit (v_log_f32; v_sin_f32 / v_cos_f32 take revolutions;
// v_sqrt_f32).  Uniforms are the top 24 bits of a word, (w >> 8) 2^-24 +
// 2^-25 in (0,1).  Block 0: the rotation axis (phi from word 0, cos(theta)
// from word 1, renormalised in FP64 so the rotation stays orthogonal to
// FP64 precision) and the uniforms of the first two Gamma candidates
// (words 2, 3, as 32-bit words: u = (w + 0.5) 2^-32); block 1: their
// normals, the cosine and sine branches of one Box-Muller pair (words 0,
// 1).  Candidates from the third on (rare) come from the FP64 form's
// stream (gamma_tail).  The Gamma acceptance test is FP32 too
// (philox_gamma_f32); the thermostat and rotation after it are FP64.
constexpr uint32_t kTagF32Cell = 0x46334330u;  // 'F3C0'

__device__ __forceinline__ float u24_open01(uint32_t w)
{
    return (float)(w >> 8) * (1.0f / 16777216.0f) + (1.0f / 33554432.0f);
}

__device__ __forceinline__ U32x4 cell_block_f32(uint64_t seed, uint64_t step, uint64_t cell, uint32_t j)
{
    return philox4x32((uint32_t)cell, (uint32_t)step, (uint32_t)(step >> 32), kTagF32Cell + j, (uint32_t)seed,
                      (uint32_t)(seed >> 32));
}

// Block 0: the axis and the candidates' uniform words.
__device__ __forceinline__ void cell_axis_f32(uint64_t seed, uint64_t step, uint64_t cell, double ax[3], uint32_t &uw0,
                                              uint32_t &uw1)
{
    const U32x4 w = cell_block_f32(seed, step, cell, 0);
    const float ph = u24_open01(w.v[0]);
    const float ct = 2.0f * u24_open01(w.v[1]) - 1.0f;
    const float st = __builtin_amdgcn_sqrtf(fmaxf(0.0f, 1.0f - ct * ct));
    const double a0 = (double)(st * __builtin_amdgcn_cosf(ph)), a1 = (double)(st * __builtin_amdgcn_sinf(ph)),
                 a2 = (double)ct;
    const double inv = 1.0 / sqrt(a0 * a0 + a1 * a1 + a2 * a2);
    ax[0] = a0 * inv;
    ax[1] = a1 * inv;
    ax[2] = a2 * inv;
    uw0 = w.v[2];
    uw1 = w.v[3];
}

// The Gamma(3(nc-1)/2, 1) variate of a cell of the FP32-bulk form from its
// first two candidates: the acceptance test of mpcd.cpp:13-43 in FP32 on
// the transcendental unit (ln as log2 times ln 2, 1/sqrt(d) by v_rsq_f32);
// the rare third candidate and later ones take the FP64 path.
__device__ __forceinline__ bool gamma_accept_f32(float d, float c, float x, float u, float &g)
{
    float v = 1.0f + c * x;
    if (v <= 0.0f)
        return false;
    v = v * v * v;
    const float x2 = x * x;
    if (u < 1.0f - 0.0331f * (x2 * x2) ||
        __builtin_amdgcn_logf(u) * 0.69314718055994531f <
            0.5f * x2 + d * ((1.0f - v) + __builtin_amdgcn_logf(v) * 0.69314718055994531f)) {
        g = d * v;
        return true;
    }
    return false;
}

__device__ __forceinline__ double philox_gamma_f32(uint64_t seed, uint64_t step, uint64_t cell, int nc, float x0,
                                                   uint32_t uw0, float x1, uint32_t uw1)
{
    const float d = 1.5f * (float)nc - 1.8333333333333333f;
    const float c = 0.33333333333333333f * __builtin_amdgcn_rsqf(d);
    float g;
    if (gamma_accept_f32(d, c, x0, (float)u32_open01(uw0), g) || gamma_accept_f32(d, c, x1, (float)u32_open01(uw1), g))
        return (double)g;
    const double dd = 1.5 * nc - 1.833333333333333333333333333333333333333333;
    return gamma_tail(seed, step, cell, dd, 0.333333333333333333333333333333333333333333 / sqrt(dd));
}

// Block 1: the candidates' normals.
__device__ __forceinline__ void cell_normals_f32(uint64_t seed, uint64_t step, uint64_t cell, float &x0, float &x1)
{
    const U32x4 w = cell_block_f32(seed, step, cell, 1);
    const float r = __builtin_amdgcn_sqrtf(-2.0f * (__builtin_amdgcn_logf(u24_open01(w.v[0])) * 0.69314718055994531f));
    const float ph = u24_open01(w.v[1]);
    x0 = r * __builtin_amdgcn_cosf(ph);
    x1 = r * __builtin_amdgcn_sinf(ph);
}

// Normals of a wall cell's virtual momentum (src/libcilia/mpcd.cpp:181-183):
// words (0,1), (2,3), (4,5) of counter {cell, step, 0, 'VIRT'}.
constexpr uint64_t kTagVirt = 0x56495254ull;

__device__ __forceinline__ void virt_normals(uint64_t seed, uint64_t step, uint64_t cell, double n[3])
{
    const U64x4 o = philox4x64(cell, step, 0, kTagVirt, seed, kKeyHi);
    n[0] = box_muller((uint32_t)o.v[0], (uint32_t)(o.v[0] >> 32));
    n[1] = box_muller((uint32_t)o.v[1], (uint32_t)(o.v[1] >> 32));
    n[2] = box_muller((uint32_t)o.v[2], (uint32_t)(o.v[2] >> 32));
}

enum : int { kWallsNone = 0, kWallsNoSlip = 1, kWallsSlip = 2 };

// Bounce-back at the planar walls z = 0 and z = ztop after a drift, in the
// operation order of MPCD::move (src/libcilia/mpcd.cpp:346-377): above the
// top wall NOSLIP reflects about the wall velocity (vwx, vwy) and goes back
// along the path for the time t2 spent past the wall, SLIP reflects z
// specularly; below the bottom wall the whole velocity is reversed.
__device__ __forceinline__ void wall_bounce(int walls, double ztop, double vwx, double vwy, double &x, double &y,
                                            double &z, double &vx, double &vy, double &vz)
{
    if (z > ztop) {
        if (walls == kWallsNoSlip) {
            const double t2 = (z - ztop) / vz;
            vx = 2 * vwx - vx;
            vy = 2 * vwy - vy;
            vz *= (-1);
            x += 2.0 * t2 * (vx - vwx);
            y += 2.0 * t2 * (vy - vwy);
            z += 2.0 * t2 * vz;
        } else {
            vz *= (-1.0);
            z = -z + ztop * 2.0;
        }
    }
    if (z < 0) {
        const double t2 = (z / vz);
        vx *= (-1.0);
        vy *= (-1.0);
        vz *= (-1.0);
        x += vx * 2.0 * t2;
        y += vy * 2.0 * t2;
        z += vz * 2.0 * t2;
    }
}

// Rodrigues matrix of src/libcilia/mpcd.cpp:285-293 (row-major).
__device__ __forceinline__ void rotmat(const double ax[3], double c, double s, double R[9])
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

// One axis of the cell index of src/libcilia/mpcd.cpp:130-138 (anull = 1):
// floor(s) - floor(s / L) * L, wrapped into [0, L).  That is floor(s) mod L
// whatever floor(s / L) rounds to, as long as it is within one of the true
// quotient (floor(s / L) = floor(floor(s) / L), and the wrap takes back a
// quotient one off), so the quotient is taken from floor(s) times invL =
// 1 / L instead of a division: the same integer.
__device__ __forceinline__ int box_axis(double s, int L, double invL)
{
    const double f = floor(s);
    int i = (int)(f - floor(f * invL) * (double)L);
    if (i < 0)
        i += L;
    if (i >= L)
        i -= L;
    return i;
}

// box_axis with the common case first: floor(s) in [0, L) is its own cell
// (floor(s) / L < 1, so the quotient is 0); other values (an unwrapped
// position outside the first image) take box_axis.  The same integer.
__device__ __forceinline__ int box_axis_near(double s, int L, double invL)
{
    const double f = floor(s);
    if (f >= 0.0 && f < (double)L)
        return (int)f;
    return box_axis(s, L, invL);
}

// FP32-bulk velocity grid: the nearest multiple of 2^-21 while that
// multiple is below 8 in magnitude (|round(v 2^21)| < 2^24), plain FP32
// rounding above (whose values are multiples of 2^-21 as well).
constexpr double kQ = 2097152.0;
constexpr double kQinv = 1.0 / 2097152.0;
constexpr double kQExact = 16777216.0;

__device__ __forceinline__ float quant_v(double v)
{
    const double r = rint(v * kQ);
    if (fabs(r) < kQExact)
        return (float)(r * kQinv);
    return (float)v;
}

// The grid for an FP32 value, and its integer q = v 2^21 (|v| < 1024, so
// |q| < 2^31).  v 2^21 is exact in FP32, and at or above 2^24 in magnitude
// it is already an integer, so one rounding gives both branches of the
// rule: below 8 the nearest multiple of 2^-21, above it v itself.
__device__ __forceinline__ float quant_fq(float v, int &q)
{
    const float r = rintf(v * 2097152.0f);
    q = __float2int_rz(r);
    return r * (1.0f / 2097152.0f);
}

__device__ __forceinline__ long long q_of(float w) { return (long long)((double)w * kQ); }

// q = w 2^21 as a 32-bit integer, exact for grid values below 1024.
__device__ __forceinline__ int q32_of(float w) { return __float2int_rz(w * 2097152.0f); }

// Lane helpers (wave64).
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ unsigned mbcnt64(uint64_t mask)
{
    return __builtin_amdgcn_mbcnt_hi((unsigned)(mask >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)mask, 0u));
}

// Direction of a move from tile `from` to tile `to` along an axis of nt
// tiles (periodic): 0, +1, -1, or 2 for "farther"; branch-free.  With two
// tiles both neighbours are the same tile and +1 is taken.
__device__ __forceinline__ int tile_dir(int from, int to, int nt)
{
    int d = to - from;
    d += d < 0 ? nt : 0;
    const int far = d == nt - 1 ? -1 : 2;
    return d == 0 ? 0 : (d == 1 ? 1 : far);
}

__host__ __device__ __forceinline__ int dir26(int dx, int dy, int dz)
{
    const int i = (dx + 1) * 9 + (dy + 1) * 3 + (dz + 1);
    return i > 13 ? i - 1 : i;
}

}  // namespace mcm
