// Tile-fused MPCD step kernels (FP32-bulk and FP64 forms) and the small
// kernels around them (upload, overflow fix-up, download, bead binning).
#include "mpcd_kernels.hpp"

namespace mcm {

// ===================================================================== FP32
// Record (7 words, 8 with ids): offsets ox oy oz in [0,1) of the shifted
// cell, velocities on the 2^-21 grid, packed cell key, [particle id].
template <int TX, int TY, int TZ, int K, bool TRACK>
__global__ __launch_bounds__(kBlock) void k_step_f32(StepArgs A)
{
    constexpr int NC = TX * TY * TZ;
    static_assert(NC <= kBlock, "one thread per cell");
    constexpr int W = TRACK ? 8 : 7;
    constexpr int CAP = K * kBlock;
    const Geom &g = A.g;
    const int tid = threadIdx.x;
    const TileCoord tc = tile_coord(g, blockIdx.x);

    __shared__ int s_seg[28];
    __shared__ long long s_base[27];
    __shared__ int s_ocnt[27];
    __shared__ int s_cnt[NC];
    __shared__ unsigned long long s_S[NC][3];
    __shared__ unsigned long long s_S2[NC];
    __shared__ double s_u[NC][3];
    __shared__ double s_sc[NC];
    __shared__ double s_R[NC][9];
    __shared__ unsigned long long s_Sn[NC][3];
    __shared__ unsigned long long s_dkey[NC];
    __shared__ int s_dslot[NC];
    __shared__ int s_bfirst[NC];
    __shared__ int s_bn[NC];
    __shared__ unsigned long long s_rfix[NC][3];
    __shared__ unsigned long long s_red[5];  // S2 before, q2 after, qmom xyz
    __shared__ double s_dred[2];             // bead energy before, after

    for (int c = tid; c < NC; c += kBlock) {
        s_cnt[c] = 0;
        s_S[c][0] = s_S[c][1] = s_S[c][2] = 0;
        s_S2[c] = 0;
        s_Sn[c][0] = s_Sn[c][1] = s_Sn[c][2] = 0;
        s_rfix[c][0] = s_rfix[c][1] = s_rfix[c][2] = 0;
        s_dkey[c] = 0;
        s_dslot[c] = 0x7fffffff;
        s_bn[c] = 0;
        s_bfirst[c] = 0;
    }
    if (tid < 27) {
        int cnt;
        long long base;
        segment_info(g, tc, tid, A.in_home_cnt, A.in_ob_cnt, cnt, base);
        s_seg[tid + 1] = cnt;
        s_base[tid] = base;
        s_ocnt[tid] = 0;
    }
    if (tid < 5)
        s_red[tid] = 0;
    if (tid < 2)
        s_dred[tid] = 0.0;
    __syncthreads();
    if (tid == 0) {
        s_seg[0] = 0;
        for (int s = 1; s <= 27; ++s)
            s_seg[s] += s_seg[s - 1];
        if (s_seg[27] > CAP)
            atomicOr(A.flags, kFlagTileCap);
    }
    // beads of this tile: first index and count per local cell
    if (A.nb) {
        const int b0 = A.b_tile_start[tc.t], b1 = A.b_tile_start[tc.t + 1];
        for (int i = b0 + tid; i < b1; i += kBlock) {
            const int lc = A.b_lc[i];
            if (i == b0 || A.b_lc[i - 1] != lc)
                s_bfirst[lc] = i;
            atomicAdd(&s_bn[lc], 1);
        }
    }
    __syncthreads();
    const int total = min(s_seg[27], CAP);

    uint32_t rec[K][W];
    int lcs[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int j = tid + k * kBlock;
        if (j < total) {
            int s = 0;
            while (j >= s_seg[s + 1])
                ++s;
            const uint32_t *src = (s == 0 ? A.in_home : A.in_ob) + (s_base[s] + (j - s_seg[s])) * W;
            ld_words<W>(src, rec[k]);
            const uint32_t key = rec[k][6];
            const int lx = (int)(key & 1023u) - tc.ox;
            const int ly = (int)((key >> 10) & 1023u) - tc.oy;
            const int lz = (int)(key >> 20) - tc.oz;
            int lc = lx + TX * (ly + TY * lz);
            if ((unsigned)lx >= (unsigned)TX || (unsigned)ly >= (unsigned)TY || (unsigned)lz >= (unsigned)TZ) {
                atomicOr(A.flags, kFlagMisbinned);
                lc = 0;
            }
            lcs[k] = lc;
            atomicAdd(&s_cnt[lc], 1);
            const long long qx = q_of(bitsf(rec[k][3]));
            const long long qy = q_of(bitsf(rec[k][4]));
            const long long qz = q_of(bitsf(rec[k][5]));
            atomicAdd(&s_S[lc][0], (unsigned long long)qx);
            atomicAdd(&s_S[lc][1], (unsigned long long)qy);
            atomicAdd(&s_S[lc][2], (unsigned long long)qz);
            atomicAdd(&s_S2[lc], (unsigned long long)(qx * qx + qy * qy + qz * qz));
        }
    }
    __syncthreads();

    // ---- per-cell thermostat + rotation (mpcd.cpp:158-212)
    if (tid < NC) {
        const int c = tid;
        const int lx = c % TX, ly = (c / TX) % TY, lz = c / (TX * TY);
        const int cx = tc.ox + lx, cy = tc.oy + ly, cz = tc.oz + lz;
        const int nfl = s_cnt[c], nbd = s_bn[c];
        const int num = nfl + nbd;
        if (num > 0 && cx < g.L[0] && cy < g.L[1] && cz < g.L[2]) {
            double bx = 0.0, by = 0.0, bz = 0.0, ble = 0.0;
            int bm = 0;
            for (int i = s_bfirst[c]; i < s_bfirst[c] + nbd; ++i) {
                const double *v = A.b_vel4 + 4 * (long long)i;
                const double m = v[3];
                bx += v[0] * m;
                by += v[1] * m;
                bz += v[2] * m;
                bm += (int)m;
                ble += m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            }
            const int masse = nfl + bm;
            const long long S0 = (long long)s_S[c][0], S1 = (long long)s_S[c][1], S2c = (long long)s_S[c][2];
            const double ux = ((double)S0 * kQinv + bx) / masse;
            const double uy = ((double)S1 * kQinv + by) / masse;
            const double uz = ((double)S2c * kQinv + bz) / masse;
            const double le = (double)(long long)s_S2[c] * (kQinv * kQinv) + ble;
            const double rel = le - (ux * ux + uy * uy + uz * uz) * masse;
            const long long gid = cx + (long long)g.L[0] * (cy + (long long)g.L[1] * cz);
            double sc = 1.0;
            if (A.thermostat && rel > 0 && num > 1) {
                const double G = A.inj_gamma ? A.inj_gamma[gid] : philox_gamma(A.seed, A.step, (uint64_t)gid, num);
                sc = sqrt(A.kT * G / (rel * 0.5));
            } else if (!(rel > 0 && num > 1) && !(rel * rel < 1e-8)) {
                atomicAdd((unsigned long long *)&A.diag->rel_viol, 1ull);
            }
            double ax[3];
            if (A.inj_axis) {
                ax[0] = A.inj_axis[3 * gid];
                ax[1] = A.inj_axis[3 * gid + 1];
                ax[2] = A.inj_axis[3 * gid + 2];
            } else {
                philox_axis(A.seed, A.step, (uint64_t)gid, ax);
            }
            double R[9];
            rotmat(ax, A.calpha, A.salpha, R);
#pragma unroll
            for (int i = 0; i < 9; ++i)
                s_R[c][i] = R[i];
            s_u[c][0] = ux;
            s_u[c][1] = uy;
            s_u[c][2] = uz;
            s_sc[c] = sc;
            atomicAdd(&s_red[0], s_S2[c]);
            if (nbd)
                atomicAdd(&s_dred[0], ble);
        }
    }
    __syncthreads();

    // ---- per-particle rotation + thermostat scaling, onto the 2^-21 grid
    uint32_t elig = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int j = tid + k * kBlock;
        if (j < total) {
            const int c = lcs[k];
            const double u0 = s_u[c][0], u1 = s_u[c][1], u2 = s_u[c][2], sc = s_sc[c];
            const double r0 = (double)bitsf(rec[k][3]) - u0;
            const double r1 = (double)bitsf(rec[k][4]) - u1;
            const double r2 = (double)bitsf(rec[k][5]) - u2;
            const double *R = s_R[c];
            const double vn0 = u0 + (R[0] * r0 + R[1] * r1 + R[2] * r2) * sc;
            const double vn1 = u1 + (R[3] * r0 + R[4] * r1 + R[5] * r2) * sc;
            const double vn2 = u2 + (R[6] * r0 + R[7] * r1 + R[8] * r2) * sc;
            const float w0 = quant_v(vn0), w1 = quant_v(vn1), w2 = quant_v(vn2);
            const long long q0 = q_of(w0), q1 = q_of(w1), q2 = q_of(w2);
            atomicAdd(&s_Sn[c][0], (unsigned long long)q0);
            atomicAdd(&s_Sn[c][1], (unsigned long long)q1);
            atomicAdd(&s_Sn[c][2], (unsigned long long)q2);
            if (s_bn[c]) {
                atomicAdd(&s_rfix[c][0], (unsigned long long)(long long)rint((vn0 * kQ - (double)q0) * kQ));
                atomicAdd(&s_rfix[c][1], (unsigned long long)(long long)rint((vn1 * kQ - (double)q1) * kQ));
                atomicAdd(&s_rfix[c][2], (unsigned long long)(long long)rint((vn2 * kQ - (double)q2) * kQ));
            }
            rec[k][3] = fbits(w0);
            rec[k][4] = fbits(w1);
            rec[k][5] = fbits(w2);
            if (fabsf(w0) < 7.5f && fabsf(w1) < 7.5f && fabsf(w2) < 7.5f) {
                elig |= 1u << k;
                atomicMax(&s_dkey[c], ((unsigned long long)rec[k][0] << 32) | rec[k][1]);
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int j = tid + k * kBlock;
        if (j < total && (elig >> k & 1u)) {
            const int c = lcs[k];
            if ((((unsigned long long)rec[k][0] << 32) | rec[k][1]) == s_dkey[c])
                atomicMin(&s_dslot[c], j);
        }
    }
    __syncthreads();

    // ---- beads of the tile (FP64); the first bead of a cell absorbs the
    //      fluid's quantisation residual
    if (tid < NC && s_bn[tid] > 0) {
        const int c = tid;
        const double u0 = s_u[c][0], u1 = s_u[c][1], u2 = s_u[c][2], sc = s_sc[c];
        const double *R = s_R[c];
        double blen = 0.0;
        for (int i = s_bfirst[c]; i < s_bfirst[c] + s_bn[c]; ++i) {
            const double *v = A.b_vel4 + 4 * (long long)i;
            const double m = v[3];
            const double r0 = v[0] - u0, r1 = v[1] - u1, r2 = v[2] - u2;
            double n0 = u0 + (R[0] * r0 + R[1] * r1 + R[2] * r2) * sc;
            double n1 = u1 + (R[3] * r0 + R[4] * r1 + R[5] * r2) * sc;
            double n2 = u2 + (R[6] * r0 + R[7] * r1 + R[8] * r2) * sc;
            if (i == s_bfirst[c]) {
                n0 += (double)(long long)s_rfix[c][0] * (kQinv * kQinv) / m;
                n1 += (double)(long long)s_rfix[c][1] * (kQinv * kQinv) / m;
                n2 += (double)(long long)s_rfix[c][2] * (kQinv * kQinv) / m;
            }
            double *o = A.b_out4 + 4 * (long long)A.b_idx[i];
            o[0] = n0;
            o[1] = n1;
            o[2] = n2;
            o[3] = m;
            blen += m * (n0 * n0 + n1 * n1 + n2 * n2);
        }
        atomicAdd(&s_dred[1], blen);
    }

    // ---- residual, stream, rebin under the next shift, write
    long long lq2 = 0, lqx = 0, lqy = 0, lqz = 0;
    bool bad_v = false;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int j = tid + k * kBlock;
        const bool valid = j < total;
        int dest = -1;  // 0 = home, 1 + d = outbox d, 27 = overflow
        if (valid) {
            const int c = lcs[k];
            if (j == s_dslot[c] && s_bn[c] == 0) {
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const long long res = (long long)s_S[c][a] - (long long)s_Sn[c][a];
                    rec[k][3 + a] = fbits((float)((double)(q_of(bitsf(rec[k][3 + a])) + res) * kQinv));
                }
            }
            const uint32_t key = rec[k][6];
            const int cc[3] = {(int)(key & 1023u), (int)((key >> 10) & 1023u), (int)(key >> 20)};
            const int tt[3] = {tc.tx, tc.ty, tc.tz};
            int nc[3], st[3];
            bool far = false;
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float w = bitsf(rec[k][3 + a]);
                const long long q = q_of(w);
                lq2 += q * q;
                if (a == 0) lqx += q;
                if (a == 1) lqy += q;
                if (a == 2) lqz += q;
                bad_v |= !(fabsf(w) < 512.0f);
                // stream in FP32: off + w*h + (a_next - a)  (fluid.cpp:155 + mpcd.cpp:249-253)
                const float t1 = __fmul_rn(w, A.hf);
                const float t2 = __fadd_rn(bitsf(rec[k][a]), t1);
                const float p = __fadd_rn(t2, A.da[a]);
                const float fl = floorf(p);
                int n = (int)fl;
                float o = __fsub_rn(p, fl);
                if (o >= 1.0f) {
                    o = 0.0f;
                    n += 1;
                }
                rec[k][a] = fbits(o);
                int cn = cc[a] + n;
                const int L = g.L[a];
                if (cn < 0)
                    cn += L;
                else if (cn >= L)
                    cn -= L;
                if (cn < 0 || cn >= L) {
                    bad_v = true;
                    cn = ((cn % L) + L) % L;
                }
                nc[a] = cn;
                const int Ta = a == 0 ? TX : (a == 1 ? TY : TZ);
                st[a] = tile_step(tt[a], cn / Ta, g.nt[a]);
                far |= st[a] == 2;
            }
            rec[k][6] = pack_cell(nc[0], nc[1], nc[2]);
            if (TRACK && A.dbg_cell) {
                const int id = (int)rec[k][7];
                A.dbg_cell[id] = cc[0] + g.L[0] * (cc[1] + g.L[1] * cc[2]);
                A.dbg_rank[id] = -1;
            }
            dest = far ? 27 : ((st[0] | st[1] | st[2]) == 0 ? 0 : 1 + dir26(st[0], st[1], st[2]));
        }
        const int hr = wave_bump(&s_ocnt[0], dest == 0);
        if (valid) {
            uint32_t *dst = nullptr;
            if (dest == 0) {
                if (hr < g.cap_home)
                    dst = A.out_home + ((long long)tc.t * g.cap_home + hr) * W;
            } else if (dest < 27) {
                const int d = dest - 1;
                const int r = atomicAdd(&s_ocnt[dest], 1);
                if (r < g.ob_cap[d])
                    dst = A.out_ob + ((long long)tc.t * g.ob_stride + g.ob_off[d] + r) * W;
            }
            if (!dst) {
                const int slot = atomicAdd(A.ovf_cnt, 1);
                if (slot < g.cap_ovf)
                    dst = A.ovf + (long long)slot * W;
                else
                    atomicOr(A.flags, kFlagOverflowCap);
            }
            if (dst)
                st_words<W>(dst, rec[k]);
        }
    }
    if (bad_v)
        atomicOr(A.flags, kFlagVelocity);
    // block reduction of the exact fluid sums
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        lq2 += __shfl_down(lq2, off);
        lqx += __shfl_down(lqx, off);
        lqy += __shfl_down(lqy, off);
        lqz += __shfl_down(lqz, off);
    }
    if (lane_id() == 0) {
        atomicAdd(&s_red[1], (unsigned long long)lq2);
        atomicAdd(&s_red[2], (unsigned long long)lqx);
        atomicAdd(&s_red[3], (unsigned long long)lqy);
        atomicAdd(&s_red[4], (unsigned long long)lqz);
    }
    __syncthreads();
    if (tid < 27) {
        const int cnt = s_ocnt[tid];
        if (tid == 0)
            A.out_home_cnt[tc.t] = min(cnt, g.cap_home);
        else
            A.out_ob_cnt[tc.t * 26 + tid - 1] = min(cnt, g.ob_cap[tid - 1]);
    }
    if (tid == 0) {
        Diag *D = A.diag;
        atomicAdd((unsigned long long *)&D->q_diss, (unsigned long long)((long long)s_red[0] - (long long)s_red[1]));
        atomicAdd((unsigned long long *)&D->q_mom[0], s_red[2]);
        atomicAdd((unsigned long long *)&D->q_mom[1], s_red[3]);
        atomicAdd((unsigned long long *)&D->q_mom[2], s_red[4]);
        atomicAdd(&D->e2, (double)(long long)s_red[1] * (kQinv * kQinv));
        atomicAdd((unsigned long long *)&D->n, (unsigned long long)total);
        if (s_dred[0] != 0.0 || s_dred[1] != 0.0)
            atomicAdd(&D->diss, s_dred[0] - s_dred[1]);
    }
}

// ===================================================================== FP64
// Record (12 words, 14 with ids): x y z vx vy vz as doubles (positions
// unwrapped, as the reference keeps them), [id, pad].
template <int NC>
__device__ __forceinline__ void wave0_exclusive_scan(const int *in, int *out)
{
    // one wave scans NC <= 256 counts, NC/64 (rounded up) per lane
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

template <int TX, int TY, int TZ, int K, bool TRACK>
__global__ __launch_bounds__(kBlock) void k_step_f64(StepArgs A)
{
    constexpr int NC = TX * TY * TZ;
    static_assert(NC <= kBlock, "one thread per cell");
    constexpr int W = TRACK ? 14 : 12;
    constexpr int CAP = K * kBlock;
    static_assert(CAP < 65536, "16-bit slots");
    const Geom &g = A.g;
    const int tid = threadIdx.x;
    const TileCoord tc = tile_coord(g, blockIdx.x);

    __shared__ int s_seg[28];
    __shared__ long long s_base[27];
    __shared__ int s_ocnt[27];
    __shared__ int s_cnt[NC];
    __shared__ int s_start[NC];
    __shared__ double s_v[CAP][3];
    __shared__ unsigned long long s_key[CAP];
    __shared__ unsigned short s_sorted[CAP];
    __shared__ double s_u[NC][3];
    __shared__ double s_sc[NC];
    __shared__ double s_R[NC][9];
    __shared__ int s_bfirst[NC];
    __shared__ int s_bn[NC];
    __shared__ double s_dred[6];  // e before, bead e after, fluid e after, mom xyz

    for (int c = tid; c < NC; c += kBlock) {
        s_cnt[c] = 0;
        s_bn[c] = 0;
        s_bfirst[c] = 0;
    }
    if (tid < 27) {
        int cnt;
        long long base;
        segment_info(g, tc, tid, A.in_home_cnt, A.in_ob_cnt, cnt, base);
        s_seg[tid + 1] = cnt;
        s_base[tid] = base;
        s_ocnt[tid] = 0;
    }
    if (tid < 6)
        s_dred[tid] = 0.0;
    __syncthreads();
    if (tid == 0) {
        s_seg[0] = 0;
        for (int s = 1; s <= 27; ++s)
            s_seg[s] += s_seg[s - 1];
        if (s_seg[27] > CAP)
            atomicOr(A.flags, kFlagTileCap);
    }
    if (A.nb) {
        const int b0 = A.b_tile_start[tc.t], b1 = A.b_tile_start[tc.t + 1];
        for (int i = b0 + tid; i < b1; i += kBlock) {
            const int lc = A.b_lc[i];
            if (i == b0 || A.b_lc[i - 1] != lc)
                s_bfirst[lc] = i;
            atomicAdd(&s_bn[lc], 1);
        }
    }
    __syncthreads();
    const int total = min(s_seg[27], CAP);

    uint32_t rec[K][W];
    int lcs[K];
    int rnk[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int j = tid + k * kBlock;
        if (j < total) {
            int s = 0;
            while (j >= s_seg[s + 1])
                ++s;
            const uint32_t *src = (s == 0 ? A.in_home : A.in_ob) + (s_base[s] + (j - s_seg[s])) * W;
            ld_words<W>(src, rec[k]);
            const double x = djoin(rec[k][0], rec[k][1]);
            const double y = djoin(rec[k][2], rec[k][3]);
            const double z = djoin(rec[k][4], rec[k][5]);
            int lx = box_axis(x + A.a[0], g.L[0]) - tc.ox;
            int ly = box_axis(y + A.a[1], g.L[1]) - tc.oy;
            int lz = box_axis(z + A.a[2], g.L[2]) - tc.oz;
            if (lx < 0 || lx >= TX || ly < 0 || ly >= TY || lz < 0 || lz >= TZ) {
                atomicOr(A.flags, kFlagMisbinned);
                lx = min(max(lx, 0), TX - 1);
                ly = min(max(ly, 0), TY - 1);
                lz = min(max(lz, 0), TZ - 1);
            }
            const int lc = lx + TX * (ly + TY * lz);
            lcs[k] = lc;
            rnk[k] = atomicAdd(&s_cnt[lc], 1);
            s_v[j][0] = djoin(rec[k][6], rec[k][7]);
            s_v[j][1] = djoin(rec[k][8], rec[k][9]);
            s_v[j][2] = djoin(rec[k][10], rec[k][11]);
            if constexpr (TRACK)
                s_key[j] = ((unsigned long long)rec[k][12] << 16) | (unsigned)j;
            else
                s_key[j] = ((unsigned long long)rec[k][1] << 32) | rec[k][0];
        }
    }
    __syncthreads();
    if (tid < 64)
        wave0_exclusive_scan<NC>(s_cnt, s_start);
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int j = tid + k * kBlock;
        if (j < total)
            s_sorted[s_start[lcs[k]] + rnk[k]] = (unsigned short)j;
    }
    __syncthreads();

    // ---- per cell: stable order, sums in that order, thermostat, rotation
    //      (mpcd.cpp:158-236).  Order = particle id (upload order) with ids,
    //      else a content key; beads follow in bead index order.
    if (tid < NC) {
        const int c = tid;
        const int lx = c % TX, ly = (c / TX) % TY, lz = c / (TX * TY);
        const int cx = tc.ox + lx, cy = tc.oy + ly, cz = tc.oz + lz;
        const int nfl = s_cnt[c], nbd = s_bn[c];
        const int num = nfl + nbd;
        if (num > 0 && cx < g.L[0] && cy < g.L[1] && cz < g.L[2]) {
            unsigned short *lst = s_sorted + s_start[c];
            for (int i = 1; i < nfl; ++i) {
                const unsigned short sj = lst[i];
                const unsigned long long kj = s_key[sj];
                int p = i - 1;
                while (p >= 0 && (s_key[lst[p]] > kj || (s_key[lst[p]] == kj && lst[p] > sj))) {
                    lst[p + 1] = lst[p];
                    --p;
                }
                lst[p + 1] = sj;
            }
            double ux = 0.0, uy = 0.0, uz = 0.0, le = 0.0;
            int masse = 0;
            for (int i = 0; i < nfl; ++i) {
                const double *v = s_v[lst[i]];
                const double m = 1.0;
                ux += v[0] * m;
                uy += v[1] * m;
                uz += v[2] * m;
                masse += (int)m;
                le += m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            }
            for (int i = s_bfirst[c]; i < s_bfirst[c] + nbd; ++i) {
                const double *v = A.b_vel4 + 4 * (long long)i;
                const double m = v[3];
                ux += v[0] * m;
                uy += v[1] * m;
                uz += v[2] * m;
                masse += (int)m;
                le += m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            }
            ux /= masse;
            uy /= masse;
            uz /= masse;
            const double rel = le - (ux * ux + uy * uy + uz * uz) * masse;
            const long long gid = cx + (long long)g.L[0] * (cy + (long long)g.L[1] * cz);
            double sc = 1.0;
            if (A.thermostat && rel > 0 && num > 1) {
                const double G = A.inj_gamma ? A.inj_gamma[gid] : philox_gamma(A.seed, A.step, (uint64_t)gid, num);
                sc = sqrt(A.kT * G / (rel * 0.5));
            } else if (!(rel > 0 && num > 1) && !(rel * rel < 1e-8)) {
                atomicAdd((unsigned long long *)&A.diag->rel_viol, 1ull);
            }
            double ax[3];
            if (A.inj_axis) {
                ax[0] = A.inj_axis[3 * gid];
                ax[1] = A.inj_axis[3 * gid + 1];
                ax[2] = A.inj_axis[3 * gid + 2];
            } else {
                philox_axis(A.seed, A.step, (uint64_t)gid, ax);
            }
            double R[9];
            rotmat(ax, A.calpha, A.salpha, R);
#pragma unroll
            for (int i = 0; i < 9; ++i)
                s_R[c][i] = R[i];
            s_u[c][0] = ux;
            s_u[c][1] = uy;
            s_u[c][2] = uz;
            s_sc[c] = sc;
            double blen = 0.0;
            for (int i = s_bfirst[c]; i < s_bfirst[c] + nbd; ++i) {
                const double *v = A.b_vel4 + 4 * (long long)i;
                const double m = v[3];
                const double r0 = v[0] - ux, r1 = v[1] - uy, r2 = v[2] - uz;
                double *o = A.b_out4 + 4 * (long long)A.b_idx[i];
                const double n0 = ux + (R[0] * r0 + R[1] * r1 + R[2] * r2) * sc;
                const double n1 = uy + (R[3] * r0 + R[4] * r1 + R[5] * r2) * sc;
                const double n2 = uz + (R[6] * r0 + R[7] * r1 + R[8] * r2) * sc;
                o[0] = n0;
                o[1] = n1;
                o[2] = n2;
                o[3] = m;
                blen += m * (n0 * n0 + n1 * n1 + n2 * n2);
            }
            atomicAdd(&s_dred[0], le);
            atomicAdd(&s_dred[1], blen);
            if (TRACK && A.dbg_cell) {
                for (int i = 0; i < nfl; ++i) {
                    const int id = (int)(s_key[lst[i]] >> 16);
                    A.dbg_cell[id] = (int)gid;
                    A.dbg_rank[id] = i;
                }
            }
        }
    }
    __syncthreads();

    // ---- per particle: rotate, stream (mpcd.cpp:340-343), rebin, write
    double le_after = 0.0, mx = 0.0, my = 0.0, mz = 0.0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const int j = tid + k * kBlock;
        const bool valid = j < total;
        int dest = -1;
        if (valid) {
            const int c = lcs[k];
            const double ux = s_u[c][0], uy = s_u[c][1], uz = s_u[c][2], sc = s_sc[c];
            const double *R = s_R[c];
            const double v0 = djoin(rec[k][6], rec[k][7]);
            const double v1 = djoin(rec[k][8], rec[k][9]);
            const double v2 = djoin(rec[k][10], rec[k][11]);
            const double r0 = v0 - ux, r1 = v1 - uy, r2 = v2 - uz;
            const double n0 = ux + (R[0] * r0 + R[1] * r1 + R[2] * r2) * sc;
            const double n1 = uy + (R[3] * r0 + R[4] * r1 + R[5] * r2) * sc;
            const double n2 = uz + (R[6] * r0 + R[7] * r1 + R[8] * r2) * sc;
            le_after += 1.0 * (n0 * n0 + n1 * n1 + n2 * n2);
            mx += n0;
            my += n1;
            mz += n2;
            double x = djoin(rec[k][0], rec[k][1]);
            double y = djoin(rec[k][2], rec[k][3]);
            double z = djoin(rec[k][4], rec[k][5]);
            x += n0 * A.h;
            y += n1 * A.h;
            z += n2 * A.h;
            dsplit(x, rec[k][0], rec[k][1]);
            dsplit(y, rec[k][2], rec[k][3]);
            dsplit(z, rec[k][4], rec[k][5]);
            dsplit(n0, rec[k][6], rec[k][7]);
            dsplit(n1, rec[k][8], rec[k][9]);
            dsplit(n2, rec[k][10], rec[k][11]);
            const int c0 = box_axis(x + A.a_next[0], g.L[0]);
            const int c1 = box_axis(y + A.a_next[1], g.L[1]);
            const int c2 = box_axis(z + A.a_next[2], g.L[2]);
            const int s0 = tile_step(tc.tx, c0 / TX, g.nt[0]);
            const int s1 = tile_step(tc.ty, c1 / TY, g.nt[1]);
            const int s2 = tile_step(tc.tz, c2 / TZ, g.nt[2]);
            const bool far = s0 == 2 || s1 == 2 || s2 == 2;
            dest = far ? 27 : ((s0 | s1 | s2) == 0 ? 0 : 1 + dir26(s0, s1, s2));
        }
        const int hr = wave_bump(&s_ocnt[0], dest == 0);
        if (valid) {
            uint32_t *dst = nullptr;
            if (dest == 0) {
                if (hr < g.cap_home)
                    dst = A.out_home + ((long long)tc.t * g.cap_home + hr) * W;
            } else if (dest < 27) {
                const int d = dest - 1;
                const int r = atomicAdd(&s_ocnt[dest], 1);
                if (r < g.ob_cap[d])
                    dst = A.out_ob + ((long long)tc.t * g.ob_stride + g.ob_off[d] + r) * W;
            }
            if (!dst) {
                const int slot = atomicAdd(A.ovf_cnt, 1);
                if (slot < g.cap_ovf)
                    dst = A.ovf + (long long)slot * W;
                else
                    atomicOr(A.flags, kFlagOverflowCap);
            }
            if (dst)
                st_words<W>(dst, rec[k]);
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        le_after += __shfl_down(le_after, off);
        mx += __shfl_down(mx, off);
        my += __shfl_down(my, off);
        mz += __shfl_down(mz, off);
    }
    if (lane_id() == 0) {
        atomicAdd(&s_dred[2], le_after);
        atomicAdd(&s_dred[3], mx);
        atomicAdd(&s_dred[4], my);
        atomicAdd(&s_dred[5], mz);
    }
    __syncthreads();
    if (tid < 27) {
        const int cnt = s_ocnt[tid];
        if (tid == 0)
            A.out_home_cnt[tc.t] = min(cnt, g.cap_home);
        else
            A.out_ob_cnt[tc.t * 26 + tid - 1] = min(cnt, g.ob_cap[tid - 1]);
    }
    if (tid == 0) {
        Diag *D = A.diag;
        atomicAdd(&D->diss, s_dred[0] - s_dred[1] - s_dred[2]);
        atomicAdd(&D->e2, s_dred[2]);
        atomicAdd(&D->mom[0], s_dred[3]);
        atomicAdd(&D->mom[1], s_dred[4]);
        atomicAdd(&D->mom[2], s_dred[5]);
        atomicAdd((unsigned long long *)&D->n, (unsigned long long)total);
    }
}

}  // namespace mcm

// ------------------------------------------------------------- launchers
namespace mcm {

void launch_step_f32(const StepArgs &A, bool track, hipStream_t s)
{
    const dim3 grid(A.g.ntiles), block(kBlock);
    if (track)
        hipLaunchKernelGGL((k_step_f32<kF32TX, kF32TY, kF32TZ, kF32K, true>), grid, block, 0, s, A);
    else
        hipLaunchKernelGGL((k_step_f32<kF32TX, kF32TY, kF32TZ, kF32K, false>), grid, block, 0, s, A);
}

void launch_step_f64(const StepArgs &A, bool track, hipStream_t s)
{
    const dim3 grid(A.g.ntiles), block(kBlock);
    if (track)
        hipLaunchKernelGGL((k_step_f64<kF64TX, kF64TY, kF64TZ, kF64K, true>), grid, block, 0, s, A);
    else
        hipLaunchKernelGGL((k_step_f64<kF64TX, kF64TY, kF64TZ, kF64K, false>), grid, block, 0, s, A);
}

}  // namespace mcm
