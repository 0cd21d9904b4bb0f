/*
 * mcm_sim.h -- C ABI of the configuration-file driver.
 *
 * The reference's `cilia` executable (src/main.cpp:40-100, 118-141) reads a
 * libconfig file (name, globalt, desiredt, omega, sizex/y/z, runtime, tstep
 * and a list md_objects of Clami, Sphere, Cilia and ConstantBeatCilia
 * entries), builds the objects (Simulation::load_config / add_md_object,
 * src/simulation.cpp:266-351), creates the solvent (Fluid::FluidFromFile,
 * src/libcilia/fluid.cpp:21-48) and runs Simulation::run
 * (src/simulation.cpp:59-158).  This driver does the same on the MPCD
 * context of mpcd_abi.h: the objects become one bead model handed to
 * mpcd_md_attach, each step is mpcd_step, and the outputs are the
 * reference's <name>.xyz and <name>.san of every Clami (src/clami.cpp:207-
 * 277) and <name>.xyz of every object that writes data
 * (src/libcilia/mdobject.cpp:83-107).
 *
 * Loading (mcm_sim_load) is host work only; mcm_sim_start and later calls
 * use the GPU.  Every entry point returns MPCD_OK (0) or a negative status
 * with a message from mcm_sim_last_error().
 */
#ifndef MCM_SIM_H
#define MCM_SIM_H

#include "mpcd_abi.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mcm_sim mcm_sim;

typedef struct mcm_sim_info {
    char name[128];          /* root "name" */
    int size[3];             /* sizex, sizey, sizez */
    double globalt, desiredt, tstep, runtime, omega;
    int nobjects;            /* Simulation::md_objects (a Clami's sphere and cilia included) */
    int nbeads, nspheres, ncilia, nattach;
    long long steps;         /* steps taken since mcm_sim_start */
} mcm_sim_info;

typedef struct mcm_sim_object {
    char type[32];           /* "Clami", "Sphere", "Cilia" or "ConstantBeatCilia" */
    char name[128];
    int write_data;          /* writes <name>.xyz (and, a Clami, <name>.san) */
    int first, nbeads;       /* its beads in the bead model */
    int sphere;              /* index of its sphere, or -1 */
    int ncilia, first_cilium;/* its cilia (consecutive) */
    int first_attach, nattach;
} mcm_sim_object;

typedef struct mcm_sim_opts {
    int precision;             /* MPCD_FP64 (the reference's) or MPCD_FP32BULK */
    unsigned long long seed;   /* Philox key; 0 = drawn from the system entropy source */
    int measureinterval;       /* steps per output: SIM_MEASUREINTERVAL (src/Makefile:3), 0 = 50 */
    long long max_steps;       /* > 0: end the run after this many steps (a multiple of */
                               /*   measureinterval is taken whole); 0 = run to desiredt */
    int restart;               /* 1: continue from <name>.ckpt in the output directory when */
                               /*   it exists (the reference's .dat restart files); 0: ignore it */
    int quiet;                 /* 1: no progress lines on stdout */
    int ref_files;             /* 1: every save also writes the reference's restart files */
                               /*   (fluid_n.dat, mpcd.dat, <name>.dat of every object), all */
                               /*   digits; they are read at the start when no <name>.ckpt is */
                               /*   there (opts.restart), so a run of either continues the other */
} mcm_sim_opts;

const char *mcm_sim_last_error(void);

/* Reads and checks the file and builds every object on the host; outputs go
 * to the file's directory. */
int mcm_sim_load(const char *cfg_path, mcm_sim **out);
int mcm_sim_free(mcm_sim *s);
int mcm_sim_get_info(const mcm_sim *s, mcm_sim_info *out);
int mcm_sim_get_object(const mcm_sim *s, int i, mcm_sim_object *out);
/* The bead model as mpcd_md_attach takes it (pointers into the driver's
 * storage, valid until mcm_sim_free): initial state before mcm_sim_start. */
int mcm_sim_md_desc(const mcm_sim *s, mpcd_md_desc *out);

/* Creates the MPCD context (rho 10, kT 1, the reference's rotation
 * constants), the solvent of sizex*sizey*sizez*10 particles and attaches the
 * bead model, or restores all of it from <name>.ckpt (opts.restart).
 * Without a checkpoint, opts.restart reads what a run of the reference
 * left: mpcd.dat (the shift), fluid_n.dat (the solvent) and the objects'
 * <name>.dat (positions, velocities, headbond, rest lengths, beat state,
 * attachment springs), as the reference's constructors do
 * (src/libcilia/mpcd.cpp:86-97, fluid.cpp:21-48, mdobject.cpp:141-151,
 * src/clami.cpp:304-345); its random-number state is not carried over. */
int mcm_sim_start(mcm_sim *s, const mcm_sim_opts *opts);
/* n steps of Simulation::next_tstep (src/simulation.cpp:217-265). */
int mcm_sim_next_tstep(mcm_sim *s, int n);
/* writedata of every object (appends to <name>.xyz / <name>.san). */
int mcm_sim_writedata(mcm_sim *s);
/* Simulation::run: blocks of measureinterval steps, each followed by the
 * outputs, until desiredt (or opts.max_steps, or the runtime limit); saves
 * <name>.ckpt every 2 hours and at the end and writes the file's globalt
 * back (src/main.cpp:94-96). */
int mcm_sim_run(mcm_sim *s);
/* <name>.ckpt and the `parameter` file now. */
int mcm_sim_save(mcm_sim *s);
/* Runs several started simulations (same tstep, same precision) together as
 * one ensemble (mpcd_ensemble_*), each with its own outputs, each ending at
 * its own desiredt / max_steps / runtime. */
int mcm_sim_run_ensemble(mcm_sim *const *sims, int n);
/* The MPCD context of a started simulation (diagnostics, downloads). */
mpcd_ctx *mcm_sim_context(mcm_sim *s);

#ifdef __cplusplus
}
#endif

#endif
