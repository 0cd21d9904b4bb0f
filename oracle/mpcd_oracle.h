/*
 * CPU oracle for the MPCD step -- TEST INFRASTRUCTURE ONLY.
 *
 * A from-scratch C restatement of the reference algorithm.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it, and
 * only as the checker (or as the timed CPU baseline); the product path never
 * calls into it.
 *
 * Every function cites the reference file:line it restates (paths relative
 * to the reference repository root).  Build: oracle/build_oracle.sh
 * (gcc -O2 -ffp-contract=off: no fused multiply-add, so every FP64 product
 * and sum rounds separately, which is what an x86-64 build of the reference
 * without -mfma does).
 */
#ifndef MPCD_ORACLE_H
#define MPCD_ORACLE_H

#include <stdint.h>

struct orc_sphere_s;

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- MT19937 */
/* src/libcilia/MersenneTwister.h: MTRand(uint32) = initialize + reload
 * (:215-220, :283-314), tempering in randInt (:179-193). */
typedef struct {
    uint32_t s[624];
    int left;
    int next;
} orc_mt;

void     orc_mt_seed(orc_mt *m, uint32_t seed);
uint32_t orc_mt_u32(orc_mt *m);
double   orc_mt_rand(orc_mt *m);     /* [0,1]  MersenneTwister.h:146-147 */
double   orc_mt_dblexc(orc_mt *m);   /* (0,1)  MersenneTwister.h:158-159 */
double   orc_normalrand(orc_mt *m);  /* src/libcilia/basicf.h:52-62 */
double   orc_rgamma_mt(orc_mt *m, int nc);              /* src/libcilia/mpcd.cpp:13-43 */
void     orc_axis_mt(orc_mt *m, double ax[3]);          /* mpcd.cpp:269-284 */
void     orc_rotmat(const double ax[3], double c, double s, double R[9]); /* mpcd.cpp:285-293 */

/* ------------------------------------------------------ Philox4x64-10 */
/* Counter-based generator of the GPU path (Salmon et al., SC'11). */
void   orc_philox4x64(const uint64_t ctr[4], const uint64_t key[2], uint64_t out[4]);
void   orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_u64_to_open01(uint64_t x);
/* Per-cell production draws: Gamma(3(nc-1)/2) variate and rotation axis,
 * derived from (seed, step, cell) exactly as the GPU kernels derive them. */
double orc_philox_gamma(uint64_t seed, uint64_t step, uint64_t cell, int nc);
void   orc_philox_candidate(uint64_t seed, uint64_t step, uint64_t cell, uint64_t blk, double *x, double *u);
void   orc_philox_axis(uint64_t seed, uint64_t step, uint64_t cell, double ax[3]);
void   orc_philox_shift(uint64_t seed, uint64_t step, double a[3]);

/* ------------------------------------------------------------- geometry */
/* src/libcilia/mpcd.cpp:130-138 (anull = 1) */
int  orc_box(const int L[3], const double a[3], const double x[3]);
void orc_box_many(const int L[3], const double a[3], const double *pos3, long n, int *cell);
/* src/libcilia/mpcd.cpp:340-343 */
void orc_stream(double *pos3, const double *vel4, long n, double ts);
/* the same with the planar-wall bounce-back of mpcd.cpp:346-377 (walls as
 * in orc_collide_cfg; vw: top wall velocity) */
void orc_stream_walls(double *pos3, double *vel4, long n, double ts, int Lz, int walls, double vwx, double vwy);
void orc_move_walls(double *x3, double *v4, double ts, int Lz, int walls, double vwx, double vwy);
/* normal number k (0..2) of a cell's virtual wall momentum, production stream */
double orc_philox_virt(uint64_t seed, uint64_t step, uint64_t cell, int k);
/* Stream with the rigid sphere's bounce-back (MPCD::move plus the sphere):
 * adds the impulses to sp->dP, sp->dL; then orc_sphere_finish applies the
 * step's impulses and moves the sphere. */
void orc_stream_sphere(double *pos3, double *vel4, long n, double ts, const int L[3], struct orc_sphere_s *sp);
void orc_sphere_finish(struct orc_sphere_s *sp, double ts, const int L[3]);
/* candidate k of cell c's virtual sphere particles: offset u (3) in [0,1) of
 * the cell and three normals, production stream */
void orc_philox_sphv(uint64_t seed, uint64_t step, uint64_t cell, uint64_t k, double u[3], double nr[3]);

/* normal number k (0..2) of the MPC-AT random velocity of the particle of
 * rank r in its cell, production stream */
double orc_philox_ran(uint64_t seed, uint64_t step, uint64_t cell, uint64_t r, int k);
/* omega = I^-1 L of the MPC-AT+a correction (I: xx yy zz xy xz yz, signed off-diagonals) */
void orc_at_omega(const double I[6], const double L[3], double w[3]);

/* src/libcilia/fluid.cpp:55-98 (create): positions, Maxwell velocities,
 * mass column, centre-of-mass velocity removed. */
void orc_fluid_create(orc_mt *m, const int L[3], long n, double *pos3, double *vel4);
/* The same on the production's Philox 'FLUD' stream (k_create_fluid):
 * positions uniform over [0, ext) per axis, velocities N(0, kT). */
void orc_fluid_create_philox(uint64_t seed, const double ext[3], double kT, long n, double *pos3, double *vel4);

/* -------------------------------------------------- binning + collision */
/* Stable counting sort of particle ids by cell: per-cell order is insertion
 * order (src/libcilia/mpcd.cpp:257-267 with one thread).  start has
 * ncells+1 entries; perm has n entries; rank[i] = position of i in its cell. */
void orc_bin(const int *cell, long n, int ncells, long *start, long *perm, int *rank);

/* RNG source of the collision. */
enum { ORC_RNG_MT = 0, ORC_RNG_INJECT = 1, ORC_RNG_PHILOX = 2 };

typedef struct {
    int L[3];
    double calpha, salpha;   /* mpcd.h:74-75 */
    double kT;               /* 1 in the reference (rgamma scale) */
    int thermostat;          /* 1: MBS scaling (mpcd.cpp:204-210); 0: sc = 1 */
    int rng;                 /* ORC_RNG_* */
    orc_mt *mt;              /* ORC_RNG_MT */
    const double *inj_gamma; /* ORC_RNG_INJECT: per cell */
    const double *inj_axis;  /* ORC_RNG_INJECT: 3 per cell */
    uint64_t seed, step;     /* ORC_RNG_PHILOX */
    double *rec_gamma;       /* optional out: per cell gamma used (NaN if none) */
    double *rec_axis;        /* optional out: 3 per cell axis used */
    long rel_violations;     /* out: count of rel<=0 cells with rel^2 >= 1e-8 */
    /* planar walls at z = 0 and z = (Lz - 1) (mpcd.cpp:175-198; SLIP or NOSLIP) */
    int walls;               /* 0: none, 1: NOSLIP, 2: SLIP */
    int rho;                 /* mpcd.h:69; the virtual particles' temperature is kT */
    const double *inj_virt;  /* ORC_RNG_INJECT: 3 normals per cell */
    double *rec_virt;        /* optional out: 3 normals per cell used (NaN if none) */
    /* collision rule: ORC_SRD (the reference's), or the Andersen-thermostat
     * rule MPC-AT without / with angular-momentum conservation (absent in
     * the reference; Noguchi, Kikuchi, Gompper, EPL 78 (2007) 10005) */
    int rule;
    const double *inj_ran;   /* MPC-AT, ORC_RNG_INJECT: 3 normals per particle, fluid i at 3i, bead j at 3(nf+j) */
    double *rec_ran;         /* optional out: the normals used, same layout */
    struct orc_sphere_s *sphere; /* optional rigid sphere (SRD only) */
} orc_collide_cfg;

enum { ORC_SRD = 0, ORC_AT = 1, ORC_AT_A = 2 };

/* A rigid sphere in the solvent (absent in the reference, whose sphere is
 * a permeable bead shell): no-slip (or slip) bounce-back of the solvent in
 * the stream, and in the collision rho candidate points per cell that the
 * sphere cuts, the ones inside becoming virtual particles moving with the
 * sphere plus thermal noise (Padding and Louis, PRE 74 (2006) 031402).
 * Their momentum change in the collision and the bounce-back impulses go to
 * the sphere, integrated in FP64 at the end of the step; during the stream
 * it moves with the velocities it had at the start of the step. */
typedef struct orc_sphere_s {
    double R, Ms, Is;        /* radius, mass, moment of inertia */
    double X[3], V[3], W[3]; /* centre in [0,L), velocity, angular velocity */
    int slip;                /* 0: no-slip bounce-back, 1: specular */
    double dP[3], dL[3];     /* impulse of the step so far (collision + stream) */
    const double *inj_sph;   /* ORC_RNG_INJECT: 3 normals per candidate, (cell, k) at 3 (rho cell + k) */
} orc_sphere;

/* One MPCD collision over nf fluid particles followed by nb beads (bead
 * velocities carry their mass in column 3), binned with shift a.
 * src/libcilia/mpcd.cpp:140-247.  Returns the dissipated energy
 * sum_cells (E_before - E_after) accumulated in long double (mpcd.h:39). */
long double orc_collide(orc_collide_cfg *cfg, const double a[3],
                        const double *pos3, double *vel4, long nf,
                        const double *bpos3, double *bvel4, long nb);

/* ------------------------------------------------------- FP32-bulk form */
/* Restatement of the GPU FP32-bulk encoding: cell index + FP32 offsets in
 * [0,1) of the shifted cell, FP32 velocities on the 2^-21 grid (|v| < 8). */
void orc_f32_encode(const int L[3], const double a[3], const double *pos3,
                    const double *vel4, long n, int *cell, float *off3, float *vel3);
void orc_f32_decode(const int L[3], const double a[3], const int *cell,
                    const float *off3, const float *vel3, long n, double *pos3, double *vel4);
/* One FP32-bulk step exactly as the GPU kernel computes it: per-cell exact
 * integer sums of q = v 2^21 and |q|^2, the collision as the affine map
 * v' = sc R v + (u - sc R u) with fused multiply-adds, new velocities on
 * the grid, the momentum residual of each fluid-only cell given to one
 * particle (content order, ids when given) in its stored velocity, beads
 * taking the sub-grid remainder, then stream by ts with each particle's
 * velocity before the residual and rebin with shift a_next.  State arrays
 * updated in place; ids may be NULL.  Returns the dissipated energy. */
long double orc_f32_step(orc_collide_cfg *cfg, const double a[3], const double a_next[3],
                         double ts, int *cell, float *off3, float *vel3, long n,
                         const double *bpos3, double *bvel4, long nb, const int *ids);

/* ------------------------------------------------- bead-spring MD half */
/* See md_oracle.c for the reference functions restated. */
typedef struct {
    int first; /* global index of bead 0 (the centre); beads first .. first+npart-1 */
    int npart;
    double radius, springconstant, headbond;
    const int *nei_start; /* npart + 1: CSR over local bead index (Sphere::neilist rows) */
    const int *nei;       /* local neighbour indices */
} orc_md_sphere;

typedef struct {
    int first, rodlength, num_anker, constant_beat;
    double bondlength, beatstrength, bending, min_bend, max_bend, m_speed, springconstant;
    double power_stroke_duration, recovery_stroke_duration;
    double midpnt, Tstroke;         /* state */
    double *newlength, *oldlength;  /* rodlength each, state */
} orc_md_cilium;

typedef struct {
    int bead0, bead1; /* sphere point (pos0/force0), cilium point (pos1/force1) */
    double l0;
} orc_md_attach;

typedef struct {
    int nb;
    double *pos3, *vel4, *force3;
    const unsigned char *mobile; /* NULL: every bead moves */
    int nsph;
    orc_md_sphere *sph;
    int ncil;
    orc_md_cilium *cil;
    int natt;
    const orc_md_attach *att;
    double att_k;
    double globalt, ststep, tstep;
    int walls, Lz;           /* bead bounce-back at the planar walls (MPCD::move) */
    double vwx, vwy;
} orc_md_system;

void orc_md_recalc(orc_md_system *s);       /* simulation.cpp:202-215 */
void orc_md_switch_logic(orc_md_system *s); /* cilia.cpp:241-258, ConstantBeatCilia.cpp:8-23 */
void orc_md_step(orc_md_system *s);         /* simulation.cpp:217-265, MD half */
/* Pieces of the beat, exposed for their pins: Cilia::length(i) of cilium
 * cil with its stall rule (cilia.cpp:79-121), the 0.01 refresh test
 * (cilia.cpp:206-207), Cilia::localbend(0, rodlength) (cilia.cpp:68-78). */
double orc_md_cilium_length(const orc_md_system *s, int cil, int i);
int orc_md_refresh_due(double globalt, double ststep);
double orc_md_localbend(const orc_md_system *s, int cil);

/* -------------------------------------- OpenMP bucket step (CPU baseline) */
/* The reference's threaded algorithm: per-thread per-cell pointer buckets,
 * static schedules, one MT19937 per thread (mpcd.cpp:140-267,
 * fluid.cpp:152-167, simulation.cpp:217-265 without MD objects). */
typedef struct orc_omp_sim orc_omp_sim;
orc_omp_sim *orc_omp_create(const int L[3], long n, double *pos3, double *vel4,
                            int nthreads, uint32_t seed);
void orc_omp_steps(orc_omp_sim *s, int steps, double tstep);
void orc_omp_destroy(orc_omp_sim *s);
int  orc_omp_threads(orc_omp_sim *s);

#ifdef __cplusplus
}
#endif
#endif
