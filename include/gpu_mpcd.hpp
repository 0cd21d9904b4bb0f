/*
 * gpu_mpcd.hpp -- the reference's MPCD and Fluid classes on the gfx950
 * library, with their method names and call order, so that Simulation,
 * MDObject and Clami keep their call sites when the bead MD stays on the
 * host.  Compiled into libmcm_mpcd.so (multi-ciliated-microswimmers_amd/
 * sim/gpu_mpcd.cpp) over the C ABI of mpcd_abi.h.
 *
 *   class MPCD   src/libcilia/mpcd.h:19-40
 *   class Fluid  src/libcilia/fluid.h:16-40
 *
 * The solvent lives on the device.  MPCD::next_tstep collides every cell
 * (solvent and the beads handed in by add_impuls since the last call),
 * writes the beads' new velocities back through the pointers add_impuls
 * retained (as the reference's collision rewrites them in place,
 * mpcd.cpp:214-236), and streams and rebins the solvent with the shift the
 * following shift_boxes() reports (the step's stream is counter-based, so
 * that shift is known ahead); Fluid::move therefore finds its work done.
 * MPCD::move is host arithmetic on one particle (the beads' drift),
 * MPCD::move of mpcd.cpp:340-377 with the planar walls of the options.
 *
 * Errors throw mcm::GpuError (the reference throws const char* or asserts).
 */
#pragma once

#include <fstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "mpcd_abi.h"

namespace mcm {

struct GpuError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

/* What the reference fixes at compile time or reads from globals (mpcd.h:67-77,
 * global.cpp), as run-time options.  tstep is World::tstep, the length of
 * the step Fluid::move streams the solvent by. */
struct MPCDOptions {
    int precision = MPCD_FP64;
    unsigned long long seed = 1;
    double tstep = 0.05;
    int rho = 10;
    double kT = 1.0;
    double calpha = -0.6427876, salpha = 0.76604444;  /* mpcd.h:74-75 */
    int thermostat = 1;
    int walls = MPCD_WALLS_NONE;
    double wall_v[2] = {0.0, 0.0};
    int collision = MPCD_SRD;
    int track_ids = 1;
    /* the reference reads the shift from ./mpcd.dat when it exists
     * (mpcd.cpp:86-97) and writes it there in its destructor (mpcd.cpp:396-402) */
    std::string shift_file = "mpcd.dat";
    bool persist_shift = true;
};

class MPCD {
  public:
    MPCD(int sizex, int sizey, int sizez);
    MPCD(int sizex, int sizey, int sizez, const MPCDOptions &opt);
    ~MPCD();
    MPCD(const MPCD &) = delete;
    MPCD &operator=(const MPCD &) = delete;

    /* MPCD::debug (mpcd.cpp:404-467): momentum per mass, particle count,
     * mass, kT and the dissipated energy of the solvent and the beads handed
     * in, printed, and one line appended to mpcd.san. */
    void debug(double globalt = 0.0);
    /* MPCD::move (mpcd.cpp:340-377): x += v ts, then the wall bounce. */
    void move(double *position, double *velocity, double ts) const;
    /* MPCD::add_impuls (mpcd.cpp:257-267): the bead joins the next
     * collision; velocity[0..2] is rewritten by next_tstep, velocity[3] is
     * its mass.  Binned with the current shift, as there. */
    void add_impuls(const double *position, double *velocity);
    /* MPCD::next_tstep (mpcd.cpp:140-247) and the solvent's stream and
     * rebin (fluid.cpp:152-167). */
    void next_tstep();
    /* MPCD::shift_boxes (mpcd.cpp:249-253): the shift of the next
     * collision becomes current. */
    void shift_boxes();

    /* mpcd.h:39: the energy the collisions took out since the caller last
     * reset it (Clami::writedata sets it to 0, clami.cpp:281-282). */
    long double diss_energy_total = 0.0;

    /* beyond the reference */
    mpcd_ctx *context() const { return ctx_; }
    const double *shift() const { return a_; }
    const int *size() const { return size_; }
    int rho() const { return opt_.rho; }
    const MPCDOptions &options() const { return opt_; }
    long long steps() const { return step_; }

  private:
    void init(int sizex, int sizey, int sizez);
    void sync_diss();
    mpcd_ctx *ctx_ = nullptr;
    MPCDOptions opt_;
    int size_[3] = {0, 0, 0};
    double a_[3] = {0.0, 0.0, 0.0};  /* the shift of the next collision */
    bool pending_shift_ = false;     /* next_tstep ran; shift_boxes not yet */
    double a_next_[3] = {0.0, 0.0, 0.0};
    long long step_ = 0;
    long double diss_seen_ = 0.0;    /* device total already counted */
    std::vector<double> bpos_, bvel_;
    std::vector<double *> bptr_;
    std::vector<double> bout_;
};

class Fluid {
  public:
    Fluid(int numberofparticles, MPCD &mpcd, std::string filename);
    /* Fluid::FluidFromFile (fluid.cpp:21-48): load the file, or create
     * sizex * sizey * sizez * rho particles when it does not exist. */
    static Fluid *FluidFromFile(std::string filename, MPCD &mpcd);
    /* Fluid::move (fluid.cpp:152-167): streamed and binned by the device in
     * MPCD::next_tstep already. */
    void move();
    /* Fluid::create (fluid.cpp:55-98) on the device. */
    void create();
    /* Fluid::load (fluid.cpp:121-150): positions, "#velocities", velocities. */
    void load(std::istream &in);
    /* Fluid::writedata (fluid.cpp:169-180 writes nothing). */
    void writedata() {}
    /* Fluid::save (fluid.cpp:100-108) with every digit (the reference's
     * std::scientific keeps 7). */
    void save();
    /* Fluid::check (fluid.cpp:110-119): the device's capacity and
     * velocity flags instead of a print of every velocity. */
    void check();
    int count() const { return numberofparticles_; }

  private:
    MPCD &mpcd_;
    int numberofparticles_;
    std::string filename_;
};

}  // namespace mcm
