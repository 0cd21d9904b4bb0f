/*
 * mpcd_abi.h -- C ABI of the MI355X (gfx950) MPCD step.
 *
 * Drop-in boundary for the solvent side of the reference's step: the C++
 * class MPCD (src/libcilia/mpcd.h:19-40) and the solvent store Fluid
 * (src/libcilia/fluid.h:16-40).  Plain pointers and sizes only; every entry
 * point returns MPCD_OK (0) or a negative status, with a message available
 * from mpcd_last_error().  No exception crosses the boundary.
 *
 * Threading: one host thread per context.  All device work of a context is
 * ordered on the context's own HIP stream; entry points that hand data back
 * to the host synchronise that stream.
 *
 * Host array layouts follow the reference's Matrix rows: positions are
 * N x 3 doubles, velocities N x 4 doubles with the particle mass in column 3
 * (src/libcilia/fluid.cpp:12-19, mdobject.h:67-70).
 */
#ifndef MPCD_ABI_H
#define MPCD_ABI_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    MPCD_OK = 0,
    MPCD_ERR_ARG = -1,      /* invalid argument or call order */
    MPCD_ERR_HIP = -2,      /* HIP runtime error */
    MPCD_ERR_CAPACITY = -3, /* a device buffer overflowed: results invalid */
    MPCD_ERR_STATE = -4     /* context not ready for this call */
};

enum { MPCD_FP64 = 0, MPCD_FP32BULK = 1 };

/* Planar walls at z = 0 and z = size[2] - 1 (the reference's compile-time
 * SLIP / NOSLIP, mpcd.cpp:175-198, 346-377). */
enum { MPCD_WALLS_NONE = 0, MPCD_WALLS_NOSLIP = 1, MPCD_WALLS_SLIP = 2 };

/* Collision rule.  MPCD_SRD is the reference's (mpcd.cpp:140-247).  The
 * Andersen-thermostat rules MPC-AT and MPC-AT+a (with angular-momentum
 * conservation) are not in the reference (Noguchi, Kikuchi, Gompper, EPL
 * 78 (2007) 10005): v' = u + xi_i - (sum m xi)/M [+ omega x r_i], xi_i =
 * sqrt(kT/m_i) N(0,1)^3, r_i the particle's offset in its shifted cell
 * (rounded to FP32) minus the cell's centre of mass, omega = I^-1 sum m
 * r x (v - xi).  No walls; tile variant 1.  In MPCD_FP32BULK the fluid's
 * xi_i lie on the 2^-21 velocity grid and its lever arms on the 2^-24 grid
 * of the cell offsets, so every fluid sum of a cell is an exact integer
 * (the result does not depend on the order of the cell's particles); the
 * residual owner keeps the fluid momentum exact as in SRD.  Injected
 * normals (mpcd_draws.ran) need track_ids = 1 in both forms. */
enum { MPCD_SRD = 0, MPCD_AT = 1, MPCD_AT_A = 2 };

typedef struct mpcd_ctx mpcd_ctx;

typedef struct mpcd_params {
    int size[3];         /* cells per axis: MPCD(int sizex,int sizey,int sizez), mpcd.h:23 */
    int rho;             /* mean solvent particles per cell (mpcd.h:69, 10)            */
    double kT;           /* temperature of the MBS thermostat (mpcd.h:68, 1)           */
    double calpha;       /* rotation constants; the reference's truncated values are   */
    double salpha;       /*   -0.6427876 and 0.76604444 (mpcd.h:74-75)                 */
    int precision;       /* MPCD_FP64 or MPCD_FP32BULK                                 */
    int thermostat;      /* 1: cell MBS thermostat (mpcd.cpp:204-210); 0: sc = 1       */
    int track_ids;       /* 1: carry particle ids (download in upload order, debug)    */
    unsigned long long seed; /* Philox key of the per-cell draws and of the shifts     */
    int tile_variant;    /* step-kernel tile shape: -1 = default, else an index into   */
                         /*   the variant table (mpcd_tile_variants)                   */
    int walls;           /* MPCD_WALLS_*: bounce-back of solvent and beads at the     */
                         /*   walls, rho - num virtual particles in the bottom and top */
                         /*   cell layers (0 = periodic in z; MPCD_FP32BULK: tile     */
                         /*   variant 0, no beads; a bounced particle's velocity goes  */
                         /*   back on its 2^-21 grid)                                  */
    double wall_v[2];    /* top wall velocity (vwx, vwy of mpcd.h:70-71), NOSLIP      */
    int collision;       /* MPCD_SRD, MPCD_AT or MPCD_AT_A                            */
} mpcd_params;

typedef struct mpcd_diag {
    double momentum[3];          /* sum m v over the solvent after the last step        */
    double twice_kinetic;        /* sum m |v|^2 over the solvent after the last step    */
    double diss_last;            /* sum over cells of (E_before - E_after) of the last  */
                                 /*   collision (l_energy, mpcd.cpp:158-246)             */
    double diss_total;           /* running total: MPCD::diss_energy_total (mpcd.h:39)  */
    long long momentum_q[3];     /* FP32-bulk: exact sum v in units of 2^-21            */
    long long n;                 /* solvent particles held                              */
    long long rel_violations;    /* cells with rel<=0 and rel^2 >= 1e-8 (mpcd.cpp:208)  */
    unsigned int flags;          /* device error bits (0 = healthy)                     */
    long long deferred_tiles;    /* tiles of the last step above the step kernel's      */
                                 /*   capacity, stepped by the larger-capacity pass      */
    long long max_tile_load;     /* most particles one tile held in the last step       */
    long long tile_capacity;     /* particles per tile the step kernel holds; the       */
                                 /*   deferred pass holds 32 per cell of the tile        */
} mpcd_diag;

/* Library version string. */
const char *mpcd_version(void);
/* Message of the last failing call on this thread. */
const char *mpcd_last_error(void);

/* Creates a context: MPCD::MPCD(sizex, sizey, sizez) (mpcd.cpp:45-98).  The
 * initial grid shift is drawn from the Philox stream (step 0) unless set with
 * mpcd_set_shift before the fluid is uploaded. */
int mpcd_create(const mpcd_params *p, mpcd_ctx **out);
int mpcd_destroy(mpcd_ctx *ctx);

/* Current random grid shift a (mpcd.h:56).  mpcd_set_shift injects the shift
 * used to bin the uploaded fluid (replaces reading mpcd.dat, mpcd.cpp:86-97);
 * it is only accepted before mpcd_upload_fluid. */
int mpcd_set_shift(mpcd_ctx *ctx, const double a[3]);
int mpcd_get_shift(mpcd_ctx *ctx, double a[3]);
/* The shift the Philox stream gives for the next step (does not commit it):
 * the replacement of MPCD::shift_boxes (mpcd.cpp:249-253). */
int mpcd_draw_shift(mpcd_ctx *ctx, double a[3]);

/* Solvent upload: Fluid::create / Fluid::load followed by add_impuls of every
 * particle (fluid.cpp:55-98, 121-150).  Column 3 of vel4 must be 1.0. */
int mpcd_upload_fluid(mpcd_ctx *ctx, const double *pos3, const double *vel4, long long n);
/* Solvent download into N x 3 / N x 4 arrays (Fluid::save content,
 * fluid.cpp:100-108).  Upload order when track_ids = 1; otherwise an
 * unspecified order.  FP32-bulk positions come back wrapped into the box. */
int mpcd_download_fluid(mpcd_ctx *ctx, double *pos3, double *vel4, long long n);
long long mpcd_count(mpcd_ctx *ctx);
/* Device-side Fluid::create (fluid.cpp:55-98): n solvent particles with
 * uniform positions and Maxwell-Boltzmann velocities at kT drawn from the
 * Philox stream, mass 1, mean velocity removed, binned with the current
 * shift.  Particle i takes Philox4x64-10 blocks {i, 0, 0, 'FLUD'} and
 * {i, 1, 0, 'FLUD'} keyed by the seed.  Positions cover the whole box with
 * walls too, as the reference's SLIP / NOSLIP build creates them
 * (fluid.cpp:59-62; its drop of fluid.cpp:67-72 needs BOUNDS, which that
 * build excludes, mpcd.cpp:389-392): the first step bounces the particles
 * above the top wall back in, so all n then lie between the walls.  No host
 * arrays are involved. */
int mpcd_create_fluid(mpcd_ctx *ctx, long long n);
/* mpcd_create_fluid that allocates nothing (no counterpart in the reference;
 * for callers that cannot allocate device memory: code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  The same fluid,
 * bit for bit: every particle is drawn twice from its counter (once for the
 * mean velocity, once to write its record straight into the home segments)
 * instead of being staged in n-sized arrays.  Its scratch is one fixed
 * 49,184-byte array per GPU (the blocks' partial sums), kept for the life of
 * the process and shared by every call on that GPU: calls that overlap on one
 * GPU race.  n beyond the records the home segments hold is refused. */
int mpcd_create_fluid_noalloc(mpcd_ctx *ctx, long long n);

/* Beads for the next collision: MDObject::md_loop_stop / init add_impuls of
 * every bead (mdobject.cpp:62-72, 153-160), binned with the current shift.
 * Column 3 of vel4 is the bead mass. */
int mpcd_stage_beads(mpcd_ctx *ctx, const double *pos3, const double *vel4, int nb);
/* Bead velocities after the last collision (the in-place update the reference
 * makes through the pointers add_impuls retained), N x 4. */
int mpcd_fetch_beads(mpcd_ctx *ctx, double *vel4, int nb);

/* One MPCD step of the solvent: collision of every cell binned with the
 * current shift (MPCD::next_tstep, mpcd.cpp:140-247, staged beads included),
 * then streaming by tstep and binning with next_shift (Fluid::move,
 * fluid.cpp:152-167, after MPCD::shift_boxes).  next_shift = NULL takes
 * mpcd_draw_shift's value.  inj_gamma (1 per cell) and inj_axis (3 per cell),
 * host arrays indexed by the reference's cell id (mpcd.cpp:130-138), replace
 * the Philox draws when non-NULL (both or neither). */
int mpcd_step(mpcd_ctx *ctx, double tstep, const double *next_shift,
              const double *inj_gamma, const double *inj_axis);
/* Host arrays of draws that replace the Philox stream in one step, for
 * bit-exact comparisons (NULL: drawn as usual).
 *   gamma, axis  SRD: 1 Gamma variate and 3 axis components per cell, indexed
 *                by the reference's cell id (mpcd.cpp:130-138), both or neither
 *   virt         walls: the three normalrand() draws of each wall cell's
 *                virtual particles (mpcd.cpp:181-183), 3 per cell; required
 *                with gamma and axis on a box with walls
 *   ran          MPC-AT: 3 normals per particle, solvent id i at 3i, bead j
 *                at 3(count + j); needs track_ids = 1; no gamma/axis/virt
 *   sph          rigid sphere: 3 normals per virtual-particle candidate,
 *                candidate k of cell c at 3(rho c + k); with gamma and axis */
typedef struct mpcd_draws {
    const double *gamma, *axis, *virt, *ran, *sph;
} mpcd_draws;

/* mpcd_step with injected draws (mpcd_step refuses injection on a box with
 * walls, a sphere or MPC-AT). */
int mpcd_step_injected(mpcd_ctx *ctx, double tstep, const double *next_shift, const mpcd_draws *draws);

/* ------------------------------------------------------ rigid sphere
 * A rigid sphere in the solvent (SRD, no walls; the FP32-bulk form with
 * tile variant 0 and no beads).  Not in the
 * reference, whose sphere is a permeable bead shell: bounce-back of the
 * solvent in the stream (no-slip: v' = 2 u_s - v at the crossing point, u_s
 * the surface velocity; slip: the normal relative velocity flips) and, in
 * the collision, rho candidate points per occupied cell that the sphere
 * cuts, the ones inside taking part as virtual unit masses moving with the
 * sphere plus thermal noise at kT.  The bounce impulses and the virtual
 * particles' momentum change move the sphere (FP64), after the stream, in
 * which it moves with the velocity it had at the start of the step.  In the
 * FP32-bulk form a bounced particle's velocity goes back on the 2^-21 grid
 * and the impulse is taken from the grid velocities, and in a cell the
 * sphere cuts the sphere takes the fluid's residual sum (q - q'), so fluid
 * and sphere momentum is conserved to FP64 roundoff. */
typedef struct mpcd_sphere {
    double radius, mass, inertia; /* inertia <= 0: 2/5 mass radius^2 */
    double x[3], v[3], w[3];      /* centre, velocity, angular velocity */
    int slip;                     /* 0: no-slip bounce-back, 1: slip */
} mpcd_sphere;

int mpcd_sphere_attach(mpcd_ctx *ctx, const mpcd_sphere *s);
int mpcd_sphere_detach(mpcd_ctx *ctx);
/* State after the last step; force and torque: the last step's impulse on
 * the sphere divided by tstep (either may be NULL). */
int mpcd_sphere_state(mpcd_ctx *ctx, mpcd_sphere *out, double force[3], double torque[3]);

/* ---------------------------------------------------------- checkpoints
 * The context's whole state between steps: the solvent records as stored
 * (exact in both forms), the grid shift, the step counter and seed that key
 * the random streams, the running dissipation totals, an attached rigid
 * sphere and an attached swimmer's dynamic state (positions, velocities,
 * last forces, beat state, rest lengths, globalt).  Loading it into a
 * context created with the same parameters (the same sphere or swimmer
 * attached, in any state) continues the run bit for bit.  Replaces the
 * reference's restart files (fluid_n.dat, <name>.dat, mpcd.dat:
 * fluid.cpp:100-150, mdobject.cpp:110-151, cilia.cpp:47-66,
 * mpcd.cpp:86-97,396-403), which store 6 significant digits. */
int mpcd_checkpoint_save(mpcd_ctx *ctx, const char *path);
int mpcd_checkpoint_load(mpcd_ctx *ctx, const char *path);

/* ------------------------------------------------------------ ensembles
 * Many independent boxes (each its own context: size, seed, fluid and
 * optional swimmer) stepped together on one GPU.  The reference runs one
 * simulation per process, whose time loop (src/simulation.cpp:217-265,
 * MPCD::step in mpcd.cpp) each box's step reproduces as mpcd_step does,
 * bit-identically.  One ensemble step is a fixed launch sequence for all
 * boxes together (per-box set-up with the swimmer's bead binning, the step
 * kernel over every box's tiles and its deferred pass, the overflow
 * fix-up, the MD half), replayed from a captured HIP graph.
 *
 * Members share precision, track_ids, tile variant and collision rule
 * (SRD, MPC-AT or MPC-AT+a), all have a rigid sphere or none has (FP32-bulk:
 * no planar walls, and no swimmer beside a sphere), have their fluid
 * loaded, no staged beads and swimmers of at most 16384 beads.  While a context is in an ensemble, calls that change
 * its state (step, upload, create_fluid, stage_beads, md_attach/detach,
 * md_run, sphere_attach/detach, destroy) are refused; reading calls
 * (download, diagnostics, md_fetch, sphere_state) work after
 * mpcd_ensemble_sync.  Every box draws its shifts and collision variates
 * from its own seed's Philox streams, as alone. */
typedef struct mpcd_ensemble mpcd_ensemble;

int mpcd_ensemble_create(mpcd_ctx *const *boxes, int nbox, mpcd_ensemble **out);
/* nsteps steps of length tstep of every box (the swimmers' MD included). */
int mpcd_ensemble_step(mpcd_ensemble *e, double tstep, int nsteps);
/* on = 0: plain launches instead of graph replays (the same kernels). */
int mpcd_ensemble_set_graph(mpcd_ensemble *e, int on);
/* The swimmers' MD half: on = 1 (default) each phase of each force
 * evaluation is one launch over every swimmer's items; on = 0 one
 * workgroup runs a swimmer's whole MD half.  Same results. */
int mpcd_ensemble_set_md_wide(mpcd_ensemble *e, int on);
/* Threads per workgroup of the wide MD launches: 256, 512 (default) or
 * 1024.  Same results. */
int mpcd_ensemble_set_md_block(mpcd_ensemble *e, int threads);
/* The ensemble's autotuner: times the MD half's launch shapes (the wide
 * path at 256, 512 and 1024 threads per workgroup, then a workgroup per
 * swimmer) on this ensemble's own steps of length tstep -- one warm-up
 * step and `steps` timed steps each, HIP events on the ensemble's stream --
 * and keeps the fastest.  The steps taken are steps of the run (every shape
 * gives the same bits): 4 (steps + 1) of them.  Out: the chosen md_wide and
 * md_block, its ms per step, and (ms_each, 4 entries, may be null) each
 * candidate's.  The step kernels' tile shape is chosen per context at
 * creation (mpcd_autotune). */
int mpcd_ensemble_autotune(mpcd_ensemble *e, double tstep, int steps, int *md_wide, int *md_block, double *ms,
                           double *ms_each);
int mpcd_ensemble_sync(mpcd_ensemble *e);
/* Releases the boxes (their contexts stay valid and usable alone). */
int mpcd_ensemble_destroy(mpcd_ensemble *e);

/* Diagnostics of the last step (synchronises): MPCD::debug (mpcd.cpp:404-467). */
int mpcd_diagnostics(mpcd_ctx *ctx, mpcd_diag *out);
/* Resets diss_total (Clami::writedata does diss_energy_total = 0, clami.cpp:282). */
int mpcd_reset_diss(mpcd_ctx *ctx);

/* Debug channel (track_ids = 1 only): when armed, each step records, per
 * particle id, the reference cell id it collided in and its position in that
 * cell's particle order; mpcd_debug_binning copies the last record out. */
int mpcd_arm_debug(mpcd_ctx *ctx, int on);
int mpcd_debug_binning(mpcd_ctx *ctx, int *cell, int *rank, long long n);

/* Tile shapes of the step kernel for a precision: writes up to max entries
 * of (tx, ty, tz, particles-per-thread) and returns the number of variants. */
int mpcd_tile_variants(int precision, int *shapes4, int max);
/* Picks the fastest tile variant for a box by timing each on a scratch
 * context with the given fluid (device-created, Philox stream): three
 * rounds of `steps` steps over the variants in turn, each variant's median
 * round; writes the winner and its step time.  The variants' contexts are
 * held together while device memory allows (one fluid each); a variant that
 * does not fit beside the others is timed in a later group.  Does not touch
 * any other context. */
int mpcd_autotune(const mpcd_params *p, long long n, int steps, int *best_variant, double *best_ms);

/* Kernel timing with HIP events on the context stream (step kernel only). */
int mpcd_set_timing(mpcd_ctx *ctx, int on);
int mpcd_timing(mpcd_ctx *ctx, double *total_ms, long long *launches);
int mpcd_sync(mpcd_ctx *ctx);

/* ------------------------------------------------------ bead MD (swimmer)
 * The MD half of the reference's step, on the device: the bead-spring
 * sub-steps of Simulation::next_tstep (src/simulation.cpp:217-265), the
 * forces of Sphere::recalcforce (src/sphere.cpp:87-119),
 * Cilia::recalcforce/addcross/length (src/cilia.cpp:80-239) and
 * Clami::global_forces (src/clami.cpp:188-198), and the beat logic
 * Cilia::switch_logic / ConstantBeatCilia::switch_logic (src/cilia.cpp:241-258,
 * src/ConstantBeatCilia.cpp:8-23).  Forces are accelerations, as there
 * (src/libcilia/Matrix.h:425-428).  Bead indices are global over one array
 * holding every object's beads (each object's beads contiguous, rod-major
 * as MDObject::positions, src/libcilia/mdobject.h:67-70). */
typedef struct mpcd_md_sphere {
    int first;            /* index of the centre bead; beads first .. first+npart-1 */
    int npart;            /* Sphere::npart */
    int nei_off;          /* this sphere's npart + 1 entries in nei_start */
    double radius, springconstant, headbond;
} mpcd_md_sphere;

typedef struct mpcd_md_cilium {
    int first;            /* bead (rod 0, bead 0); 3 x rodlength beads follow */
    int rodlength, num_anker;
    int constant_beat;    /* 1: ConstantBeatCilia, 0: Cilia */
    double bondlength, beatstrength, bending, min_bend, max_bend, m_speed, springconstant;
    double power_stroke_duration, recovery_stroke_duration;
    double midpnt, Tstroke; /* beat state */
} mpcd_md_cilium;

typedef struct mpcd_md_attachment {
    int bead0;            /* the sphere point (SpringCfg pos0/force0, src/utils.h:8-33) */
    int bead1;            /* the cilium point (pos1/force1) */
    double l0;
} mpcd_md_attachment;

typedef struct mpcd_md_desc {
    int nb;
    const double *pos3;   /* nb x 3 */
    const double *vel4;   /* nb x 4, mass in column 3 */
    const unsigned char *mobile; /* nb, 0 = fixed bead (index below startidx); NULL = all move */
    int nsph;
    const mpcd_md_sphere *sph;
    const int *nei_start; /* CSR rows (Sphere::neilist): local bead -> range in nei */
    const int *nei;       /* local neighbour indices */
    int ncil;
    const mpcd_md_cilium *cil;
    const double *newlength; /* rodlength per cilium, concatenated in cilium order */
    const double *oldlength;
    int natt;
    const mpcd_md_attachment *att; /* in Clami's order (cilium, then spring) */
    double att_springconstant;
    double globalt;       /* World::globalt at the start of the next step */
    double ststep;        /* World::ststep (src/libcilia/global.cpp:11) */
} mpcd_md_desc;

/* Hands the swimmer's beads to the context.  From then on mpcd_step
 * collides them with the solvent (binned from their device positions),
 * runs the MD sub-steps after the collision and leaves them binned for the
 * next one; mpcd_stage_beads is refused while a swimmer is attached. */
int mpcd_md_attach(mpcd_ctx *ctx, const mpcd_md_desc *desc);
int mpcd_md_detach(mpcd_ctx *ctx);
/* The MD half of one step alone (no collision): for swimmers without a
 * solvent, and for checking the sub-steps on their own. */
int mpcd_md_run(mpcd_ctx *ctx, double tstep);
/* Bead state, forces of the last evaluation, beat state (midpnt, Tstroke
 * per cilium), rest lengths and globalt; any pointer may be NULL. */
int mpcd_md_fetch(mpcd_ctx *ctx, double *pos3, double *vel4, double *force3, double *beat2, double *newlength,
                  double *oldlength, double *globalt);

#ifdef __cplusplus
}
#endif
#endif
