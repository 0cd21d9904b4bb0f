/*
 * CPU oracle for the bead-spring MD sub-steps -- TEST INFRASTRUCTURE ONLY.
 *
 * A restatement of the reference's MD half of one MPCD step, in the
 * reference's own loop order, so that each bead's force is accumulated in
 * the same order as there:
 *   MDObject::spring            src/libcilia/mdobject.cpp:8-17
 *   MDObject::spring (vector)   src/libcilia/mdobject.cpp:19-29
 *   md_loop_start/loop/stop     src/libcilia/mdobject.cpp:36-72
 *   Sphere::recalcforce         src/sphere.cpp:87-119 (REP_WALL off)
 *   Cilia::recalcforce/length/addcross/localbend/switch_logic
 *                               src/cilia.cpp:68-121, 127-258
 *   ConstantBeatCilia::switch_logic  src/ConstantBeatCilia.cpp:8-23
 *   Clami::global_forces        src/clami.cpp:188-198
 *   Simulation::next_tstep      src/simulation.cpp:217-265 (MD half)
 * Built with -ffp-contract=off.  pow(x, 2.0) and pow(x, 2.5) of the beat
 * rule are evaluated as x*x and (x*x)*sqrt(x) (what a -ffast-math build of
 * the reference computes; the GPU does the same), so every operation here
 * is a basic IEEE operation.
 */
#include <math.h>
#include <string.h>

#include "mpcd_oracle.h"

static inline double sqdist3(const double *a, const double *b)
{
    /* Matrix::sqdistance (Matrix.h:412-419): sum of squared differences in order */
    double s = 0.0;
    for (int i = 0; i < 3; ++i) {
        const double d = a[i] - b[i];
        s += d * d;
    }
    return s;
}

/* MDObject::spring(i, j, b, k): one-sided, force on bead i only. */
static void spring1(orc_md_system *s, int i, int j, double b, double k)
{
    const double *p1 = s->pos3 + 3 * (long)i, *p0 = s->pos3 + 3 * (long)j;
    double *f = s->force3 + 3 * (long)i;
    const double fac = -2.0 * k * (1.0 - b / sqrt(sqdist3(p1, p0)));
    f[0] += (p1[0] - p0[0]) * fac;
    f[1] += (p1[1] - p0[1]) * fac;
    f[2] += (p1[2] - p0[2]) * fac;
}

/* length(i) with the beat profile and the stall rule (cilia.cpp:80-121). */
static double cilium_length(const orc_md_system *s, const orc_md_cilium *c, int i)
{
    const double bl = c->bondlength;
    const double *x = s->pos3 + 3 * (long)c->first; /* rod 0 */
    double l;
    double tl = sqrt(sqdist3(x + 3 * i, x + 3 * (i + 1)));
    if (i < c->midpnt - 2.0) {
        const double t = 1.0 - (i - c->num_anker - 1.0) / (c->rodlength - c->num_anker - 1.0);
        const double p25 = (t * t) * sqrt(t); /* pow(t, 2.5) */
        l = bl + c->beatstrength * p25 * 0.5 * (1.0 - 1.0 / (c->midpnt - i - 1.0));
    } else {
        const double d = i - c->midpnt;
        l = bl - 2.0 * c->beatstrength / (d * d + 1.0); /* pow(d, 2.0) */
    }
    if (fabs(l - tl) > c->bending) {
        tl = tl + c->bending - 2.0 * c->bending * (double)(l < tl);
        if (((l - c->oldlength[i]) * (tl - c->oldlength[i])) > 0)
            l = tl;
        else
            l = c->oldlength[i];
    }
    return l;
}

/* The rest-length refresh test of Cilia::recalcforce (cilia.cpp:206-207):
 * fmod(globalt + ststep / 10, 0.01) < ststep. */
int orc_md_refresh_due(double globalt, double ststep)
{
    return fmod(globalt + ststep / 10., 0.01) < ststep;
}

double orc_md_cilium_length(const orc_md_system *s, int cil, int i)
{
    return cilium_length(s, s->cil + cil, i);
}

static void cilium_recalc(orc_md_system *s, orc_md_cilium *c)
{
    const int R = c->rodlength, o = c->first;
    const double bl = c->bondlength, k = c->springconstant;
    const double ck = 1.0 * k;                 /* crossstrength * springconstant (cilia.h:82) */
    const double crosslength = sqrt(2.0) * bl; /* cilia.h:79 */
    if (orc_md_refresh_due(s->globalt, s->ststep)) {
        for (int i = c->num_anker - 1; i < R - 1; ++i) {
            c->newlength[i] = cilium_length(s, c, i);
            c->oldlength[i] = c->newlength[i];
        }
    }
    memset(s->force3 + 3 * (long)o, 0, sizeof(double) * 9 * (size_t)R);
    for (int i = 0; i < R - 1; ++i) {
        spring1(s, o + i, o + i + 1, c->newlength[i], k);
        spring1(s, o + i + 1, o + i, c->newlength[i], k);
        spring1(s, o + R + i, o + R + i + 1, bl, k);
        spring1(s, o + R + i + 1, o + R + i, bl, k);
        spring1(s, o + 2 * R + i, o + 2 * R + i + 1, bl, k);
        spring1(s, o + 2 * R + i + 1, o + 2 * R + i, bl, k);
    }
    /* addcross (cilia.cpp:127-201) */
    for (int i = 0; i < R; ++i) {
        spring1(s, o + i, o + i + R, bl, k);
        spring1(s, o + i, o + i + 2 * R, bl, k);
        spring1(s, o + i + R, o + i, bl, k);
        spring1(s, o + i + R, o + i + 2 * R, bl, k);
        spring1(s, o + i + 2 * R, o + i, bl, k);
        spring1(s, o + i + 2 * R, o + i + R, bl, k);
    }
    for (int i = 0; i < R - 1; ++i) {
        const double li = sqrt(bl * bl + bl * c->newlength[i]);
        spring1(s, o + i, o + i + R + 1, li, ck);
        spring1(s, o + i, o + i + 2 * R + 1, li, ck);
        spring1(s, o + i + R, o + i + 1, li, ck);
        spring1(s, o + i + 2 * R, o + i + 1, li, ck);
        spring1(s, o + i + R, o + i + 2 * R + 1, crosslength, ck);
        spring1(s, o + i + 2 * R, o + i + R + 1, crosslength, ck);
    }
    for (int i = 1; i < R; ++i) {
        const double li = sqrt(bl * bl + bl * c->newlength[i - 1]);
        spring1(s, o + i, o + i + R - 1, li, ck);
        spring1(s, o + i, o + i + 2 * R - 1, li, ck);
        spring1(s, o + i + R, o + i - 1, li, ck);
        spring1(s, o + i + 2 * R, o + i - 1, li, ck);
        spring1(s, o + i + R, o + i + 2 * R - 1, crosslength, ck);
        spring1(s, o + i + 2 * R, o + i + R - 1, crosslength, ck);
    }
}

static void sphere_recalc(orc_md_system *s, const orc_md_sphere *sp)
{
    const int o = sp->first;
    for (int i = 1; i < sp->npart; i++) {
        spring1(s, o + 0, o + i, sp->radius, 0.1 * sp->springconstant);
        spring1(s, o + i, o + 0, sp->radius, 0.1 * sp->springconstant);
    }
    for (int i = 1; i < sp->npart; i++)
        for (int n = sp->nei_start[i]; n < sp->nei_start[i + 1]; ++n)
            spring1(s, o + i, o + sp->nei[n], sp->headbond, sp->springconstant);
}

/* Clami::global_forces: Vector3d spring(pos1 = sphere point, pos0 = cilium
 * point); force0 += F, force1 -= F; nothing when the points coincide. */
static void attachments(orc_md_system *s)
{
    for (int a = 0; a < s->natt; ++a) {
        const orc_md_attach *at = s->att + a;
        const double *p1 = s->pos3 + 3 * (long)at->bead0, *p0 = s->pos3 + 3 * (long)at->bead1;
        double F[3] = {0.0, 0.0, 0.0};
        const double d = sqrt(sqdist3(p1, p0));
        if (d != 0) {
            const double fac = -2.0 * s->att_k * (1.0 - at->l0 / d);
            for (int i = 0; i < 3; ++i)
                F[i] = (p1[i] - p0[i]) * fac;
        }
        double *f0 = s->force3 + 3 * (long)at->bead0, *f1 = s->force3 + 3 * (long)at->bead1;
        for (int i = 0; i < 3; ++i) {
            f0[i] += F[i];
            f1[i] -= F[i];
        }
    }
}

void orc_md_recalc(orc_md_system *s)
{
    for (int i = 0; i < s->nsph; ++i)
        sphere_recalc(s, s->sph + i);
    for (int i = 0; i < s->ncil; ++i)
        cilium_recalc(s, s->cil + i);
    attachments(s);
}

/* v += F c (Matrix::add_mul3), then x += v ts (MPCD::move) unless drift = 0;
 * forces = 0 (mdobject.cpp:36-72).  Fixed beads (mobile == 0) are skipped. */
static void md_kick_drift(orc_md_system *s, double c, int drift)
{
    for (int b = 0; b < s->nb; ++b) {
        if (s->mobile && !s->mobile[b])
            continue;
        double *x = s->pos3 + 3 * (long)b, *v = s->vel4 + 4 * (long)b, *f = s->force3 + 3 * (long)b;
        for (int i = 0; i < 3; ++i)
            v[i] += f[i] * c;
        if (drift)
            orc_move_walls(x, v, s->ststep, s->Lz, s->walls, s->vwx, s->vwy);
        f[0] = f[1] = f[2] = 0.0;
    }
}

static double localbend(const orc_md_system *s, const orc_md_cilium *c)
{
    const int R = c->rodlength;
    const double *x = s->pos3 + 3 * (long)c->first;
    double l = 0;
    for (int i = 0; i < R - 1; i++) {
        l += 2.0 * sqrt(sqdist3(x + 3 * i, x + 3 * (i + 1)));
        l -= sqrt(sqdist3(x + 3 * (R + i), x + 3 * (R + i + 1)));
        l -= sqrt(sqdist3(x + 3 * (2 * R + i), x + 3 * (2 * R + i + 1)));
    }
    return l;
}

double orc_md_localbend(const orc_md_system *s, int cil)
{
    return localbend(s, s->cil + cil);
}

void orc_md_switch_logic(orc_md_system *s)
{
    for (int i = 0; i < s->ncil; ++i) {
        orc_md_cilium *c = s->cil + i;
        const double start = c->num_anker;
        if (c->constant_beat) {
            if (s->globalt - c->Tstroke >= c->power_stroke_duration && c->midpnt == start)
                c->midpnt += 1;
            if (s->globalt - c->Tstroke - c->power_stroke_duration >= c->recovery_stroke_duration &&
                c->midpnt > start) {
                c->Tstroke = s->globalt;
                c->midpnt = start;
            }
        } else {
            if ((c->midpnt == start) && s->globalt - c->Tstroke > 100.0)
                c->midpnt += 1;
            if ((localbend(s, c) > c->max_bend) && (c->midpnt > start)) {
                c->Tstroke = s->globalt;
                c->midpnt = start;
            }
        }
        if (c->midpnt > start)
            c->midpnt += s->tstep / c->m_speed;
    }
}

/* The MD half of Simulation::next_tstep after the collision: recalc,
 * md_loop_start, recalc, (n-1) x [globalt += ststep, md_loop, recalc],
 * globalt += ststep, md_loop_stop (the half kick), switch_logic. */
void orc_md_step(orc_md_system *s)
{
    const int n = (int)((s->tstep + 0.00000001) / s->ststep);
    orc_md_recalc(s);
    md_kick_drift(s, s->ststep / 2.0, 1);
    orc_md_recalc(s);
    for (int k = 1; k < n; k++) {
        s->globalt += s->ststep;
        md_kick_drift(s, s->ststep, 1);
        orc_md_recalc(s);
    }
    s->globalt += s->ststep;
    md_kick_drift(s, s->ststep / 2.0, 0);
    orc_md_switch_logic(s);
}
