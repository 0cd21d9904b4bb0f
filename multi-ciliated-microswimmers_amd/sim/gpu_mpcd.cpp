// The reference's MPCD / Fluid classes over the C ABI (include/gpu_mpcd.hpp).
#include "gpu_mpcd.hpp"

#include <cmath>
#include <cstdio>
#include <iostream>
#include <sstream>

namespace mcm {

namespace {

void ok_or_throw(int rc, const std::string &what)
{
    if (rc != MPCD_OK)
        throw GpuError(what + ": " + mpcd_last_error());
}

bool read_shift(const std::string &path, double a[3])
{
    std::ifstream in(path);
    if (!in.good())
        return false;
    double v[3];
    if (!(in >> v[0] >> v[1] >> v[2]))
        return false;
    for (int k = 0; k < 3; ++k)
        a[k] = v[k];
    return true;
}

// The next number of a restart file, the "#tag" words skipped by the caller.
double next_number(std::istream &in, const char *what)
{
    std::string w;
    if (!(in >> w))
        throw GpuError(std::string("unexpected end of ") + what);
    char *end = nullptr;
    const double v = std::strtod(w.c_str(), &end);
    if (end == w.c_str() || *end)
        throw GpuError(std::string(what) + ": not a number: " + w);
    return v;
}

}  // namespace

// ------------------------------------------------------------------ MPCD
MPCD::MPCD(int sizex, int sizey, int sizez) { init(sizex, sizey, sizez); }

MPCD::MPCD(int sizex, int sizey, int sizez, const MPCDOptions &opt) : opt_(opt) { init(sizex, sizey, sizez); }

void MPCD::init(int sizex, int sizey, int sizez)
{
    mpcd_params p{};
    p.size[0] = sizex;
    p.size[1] = sizey;
    p.size[2] = sizez;
    p.rho = opt_.rho;
    p.kT = opt_.kT;
    p.calpha = opt_.calpha;
    p.salpha = opt_.salpha;
    p.precision = opt_.precision;
    p.thermostat = opt_.thermostat;
    p.track_ids = opt_.track_ids;
    p.seed = opt_.seed;
    p.tile_variant = -1;
    p.walls = opt_.walls;
    p.wall_v[0] = opt_.wall_v[0];
    p.wall_v[1] = opt_.wall_v[1];
    p.collision = opt_.collision;
    ok_or_throw(mpcd_create(&p, &ctx_), "MPCD");
    size_[0] = sizex;
    size_[1] = sizey;
    size_[2] = sizez;
    // the shift from mpcd.dat when it is there (mpcd.cpp:86-97), else the
    // stream's first
    double a[3];
    if (!opt_.shift_file.empty() && read_shift(opt_.shift_file, a))
        ok_or_throw(mpcd_set_shift(ctx_, a), "MPCD: the shift of mpcd.dat");
    ok_or_throw(mpcd_get_shift(ctx_, a_), "MPCD");
}

MPCD::~MPCD()
{
    if (!ctx_)
        return;
    // mpcd.cpp:396-402 meant to keep the shift for a restart (it prints the
    // array's address); the three numbers, every digit
    if (opt_.persist_shift && !opt_.shift_file.empty()) {
        if (std::FILE *f = std::fopen(opt_.shift_file.c_str(), "w")) {
            std::fprintf(f, "%.17g %.17g %.17g\n", a_[0], a_[1], a_[2]);
            std::fclose(f);
        }
    }
    mpcd_destroy(ctx_);
}

void MPCD::move(double *position, double *velocity, double ts) const
{
    position[0] += velocity[0] * ts;
    position[1] += velocity[1] * ts;
    position[2] += velocity[2] * ts;
    if (opt_.walls == MPCD_WALLS_NONE)
        return;
    // bounce-back at z = (sizez - 1) a and z = 0, in the order of
    // mpcd.cpp:346-377 (a = 1)
    const double ztop = (size_[2] - 1) * 1.0;
    if (position[2] > ztop) {
        if (opt_.walls == MPCD_WALLS_NOSLIP) {
            const double t2 = (position[2] - ztop) / velocity[2];
            velocity[0] = 2 * opt_.wall_v[0] - velocity[0];
            velocity[1] = 2 * opt_.wall_v[1] - velocity[1];
            velocity[2] *= (-1);
            position[0] += 2.0 * t2 * (velocity[0] - opt_.wall_v[0]);
            position[1] += 2.0 * t2 * (velocity[1] - opt_.wall_v[1]);
            position[2] += 2.0 * t2 * velocity[2];
        } else {
            velocity[2] *= (-1.0);
            position[2] = -position[2] + ztop * 2.0;
        }
    }
    if (position[2] < 0) {
        const double t2 = (position[2] / velocity[2]);
        velocity[0] *= (-1.0);
        velocity[1] *= (-1.0);
        velocity[2] *= (-1.0);
        position[0] += velocity[0] * 2.0 * t2;
        position[1] += velocity[1] * 2.0 * t2;
        position[2] += velocity[2] * 2.0 * t2;
    }
}

void MPCD::add_impuls(const double *position, double *velocity)
{
    bpos_.insert(bpos_.end(), position, position + 3);
    bvel_.insert(bvel_.end(), velocity, velocity + 4);
    bptr_.push_back(velocity);
}

void MPCD::sync_diss()
{
    mpcd_diag d{};
    ok_or_throw(mpcd_diagnostics(ctx_, &d), "MPCD::next_tstep");
    diss_energy_total += (long double)d.diss_total - diss_seen_;
    diss_seen_ = d.diss_total;
}

void MPCD::next_tstep()
{
    const int nb = (int)bptr_.size();
    ok_or_throw(mpcd_stage_beads(ctx_, nb ? bpos_.data() : nullptr, nb ? bvel_.data() : nullptr, nb),
          "MPCD::next_tstep (beads)");
    // the stream's next shift: the one shift_boxes() will make current
    ok_or_throw(mpcd_draw_shift(ctx_, a_next_), "MPCD::next_tstep");
    ok_or_throw(mpcd_step(ctx_, opt_.tstep, a_next_, nullptr, nullptr), "MPCD::next_tstep");
    if (nb) {
        bout_.resize(4 * (size_t)nb);
        ok_or_throw(mpcd_fetch_beads(ctx_, bout_.data(), nb), "MPCD::next_tstep (beads)");
        // the reference rewrites velocity[0..2] through the retained pointers
        for (int i = 0; i < nb; ++i)
            for (int k = 0; k < 3; ++k)
                bptr_[i][k] = bout_[4 * (size_t)i + k];
    }
    bpos_.clear();
    bvel_.clear();
    bptr_.clear();
    sync_diss();
    pending_shift_ = true;
    ++step_;
}

void MPCD::shift_boxes()
{
    // the device binned the solvent with a_next_ in next_tstep already
    if (pending_shift_) {
        for (int k = 0; k < 3; ++k)
            a_[k] = a_next_[k];
        pending_shift_ = false;
    }
}

void MPCD::debug(double globalt)
{
    mpcd_diag d{};
    ok_or_throw(mpcd_diagnostics(ctx_, &d), "MPCD::debug");
    double p[3] = {d.momentum[0], d.momentum[1], d.momentum[2]};
    double e2 = d.twice_kinetic;
    long long num = d.n;
    long long masse = d.n;  // solvent mass 1 (fluid.h:39)
    const size_t nb = bptr_.size();
    for (size_t i = 0; i < nb; ++i) {
        const double *v = &bvel_[4 * i];
        const double m = v[3];
        for (int k = 0; k < 3; ++k)
            p[k] += v[k] * m;
        masse += (long long)m;
        e2 += m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        ++num;
    }
    const double ux = p[0] / (double)masse, uy = p[1] / (double)masse, uz = p[2] / (double)masse;
    const double kT = e2 / 3.0 / (double)num;
    std::cout << "------------------------------------------------------------\n"
              << "Total Momentum \t= " << ux << " " << uy << " " << uz << "\n"
              << "Total Number \t= " << num << "\n"
              << "Total Mass \t= " << masse << "\n"
              << "kbT \t= " << kT << "\n\n"
              << "Energy diff = " << d.diss_last << "\n\n"
              << "Energy total = " << (double)diss_energy_total << "\n\n";
    std::ofstream data("mpcd.san", std::ios::out | std::ios::app);
    data << globalt << " " << (double)diss_energy_total << " " << num << " " << masse << " " << ux << " " << uy
         << " " << uz << " " << kT << "\n";
}

// ------------------------------------------------------------------ Fluid
Fluid::Fluid(int numberofparticles, MPCD &mpcd, std::string filename)
    : mpcd_(mpcd), numberofparticles_(numberofparticles), filename_(std::move(filename))
{
}

Fluid *Fluid::FluidFromFile(std::string filename, MPCD &mpcd)
{
    const int *L = mpcd.size();
    const long long want = (long long)L[0] * L[1] * L[2] * mpcd.rho();
    std::ifstream in(filename);
    if (!in.good()) {
        Fluid *fl = new Fluid((int)want, mpcd, filename);
        fl->create();
        return fl;
    }
    std::string tag;
    in >> tag;
    if (tag != "#numberofparticles")
        throw GpuError(filename + ": expected #numberofparticles");
    const long long n = (long long)next_number(in, filename.c_str());
    if (n > want)
        throw GpuError(filename + ": Fluidfile to big.");
    Fluid *fl = new Fluid((int)n, mpcd, filename);
    try {
        fl->load(in);
    } catch (...) {
        delete fl;
        throw;
    }
    return fl;
}

void Fluid::move() {}

void Fluid::create() { ok_or_throw(mpcd_create_fluid(mpcd_.context(), numberofparticles_), "Fluid::create"); }

void Fluid::load(std::istream &in)
{
    const size_t n = (size_t)numberofparticles_;
    std::vector<double> p(3 * n), v(4 * n);
    for (double &x : p)
        x = next_number(in, "fluid positions");
    std::string tag;
    in >> tag;
    if (tag != "#velocities")
        throw GpuError("fluid file: expected #velocities, found " + tag);
    for (double &x : v)
        x = next_number(in, "fluid velocities");
    ok_or_throw(mpcd_upload_fluid(mpcd_.context(), p.data(), v.data(), (long long)n), "Fluid::load");
}

void Fluid::save()
{
    const long long n = mpcd_count(mpcd_.context());
    std::vector<double> p(3 * (size_t)n), v(4 * (size_t)n);
    ok_or_throw(mpcd_download_fluid(mpcd_.context(), p.data(), v.data(), n), "Fluid::save");
    std::FILE *f = std::fopen(filename_.c_str(), "w");
    if (!f)
        throw GpuError("cannot write " + filename_);
    std::fprintf(f, "#numberofparticles %lld\n", n);
    for (long long i = 0; i < n; ++i)
        std::fprintf(f, "%.17e %.17e %.17e\n", p[3 * i], p[3 * i + 1], p[3 * i + 2]);
    std::fprintf(f, "#velocities\n");
    for (long long i = 0; i < n; ++i)
        std::fprintf(f, "%.17e %.17e %.17e %.17e\n", v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
    const bool ok = std::ferror(f) == 0;
    std::fclose(f);
    if (!ok)
        throw GpuError("cannot write " + filename_);
}

void Fluid::check()
{
    mpcd_diag d{};
    ok_or_throw(mpcd_diagnostics(mpcd_.context(), &d), std::string("Fluid::check"));
    if (d.flags)
        throw GpuError("Fluid::check: device flags " + std::to_string(d.flags));
}

}  // namespace mcm
