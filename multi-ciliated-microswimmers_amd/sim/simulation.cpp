// Configuration-file driver: the reference's Simulation (src/simulation.cpp)
// and main (src/main.cpp) over the MPCD context.  C ABI in mcm_sim.h.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <ctime>
#include <fstream>
#include <iostream>
#include <random>
#include <sstream>
#include <string>
#include <vector>

#include "cfg.hpp"
#include "mcm_sim.h"
#include "objects.hpp"

using mcm::CfgError;
using mcm::Setting;

namespace {

thread_local std::string g_err;

struct SimError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

void check(int rc, const char *what) {
    if (rc != MPCD_OK) throw SimError(std::string(what) + ": " + mpcd_last_error());
}

constexpr double kStstep = 0.001;  // World::ststep, src/libcilia/global.cpp:11
constexpr int kRho = 10;           // MPCD::rho, src/libcilia/mpcd.h:69
constexpr double kSaveEvery = 7200.0;  // seconds between saves, src/simulation.cpp:124

template <class F>
int guarded(F &&f) {
    try {
        f();
        return MPCD_OK;
    } catch (const CfgError &e) {
        g_err = e.what();
        return MPCD_ERR_ARG;
    } catch (const SimError &e) {
        g_err = e.what();
        return MPCD_ERR_STATE;
    } catch (const std::exception &e) {
        g_err = e.what();
        return MPCD_ERR_ARG;
    }
}

std::string dir_of(const std::string &path) {
    size_t k = path.find_last_of('/');
    if (k == std::string::npos) return ".";
    if (k == 0) return "/";
    return path.substr(0, k);
}

void copy_str(char *dst, size_t cap, const std::string &s) {
    std::snprintf(dst, cap, "%s", s.c_str());
}

bool file_exists(const std::string &path) { return std::ifstream(path).good(); }

// Whitespace-separated tokens of the reference's restart files (the Matrix
// text layout of src/libcilia/MatrixIO.h: rows of numbers, "#name" tags).
struct Tokens {
    std::ifstream in;
    std::string path;
    explicit Tokens(const std::string &p) : in(p), path(p) {
        if (!in.good()) throw SimError("cannot read " + p);
    }
    std::string word() {
        std::string w;
        if (!(in >> w)) throw SimError(path + ": unexpected end of file");
        return w;
    }
    void tag(const char *t) {
        const std::string w = word();
        if (w != t) throw SimError(path + ": expected " + t + ", found " + w);
    }
    double num() {
        const std::string w = word();
        char *end = nullptr;
        const double v = std::strtod(w.c_str(), &end);
        if (end == w.c_str() || *end) throw SimError(path + ": not a number: " + w);
        return v;
    }
    long long integer() {
        const double v = num();
        if (v != std::floor(v)) throw SimError(path + ": not an integer");
        return (long long)v;
    }
};

// "#name value" of MDObject::save's WRITE2 (src/libcilia/global.h:30).
long long tagged_int(Tokens &t, const char *tag) {
    t.tag(tag);
    return t.integer();
}

std::string full(double v) {
    char buf[40];
    std::snprintf(buf, sizeof buf, "%.17g", v);
    return buf;
}

}  // namespace

struct mcm_sim {
    std::string cfg_path, dir, name;
    Setting root;
    int size[3] = {0, 0, 0};
    double globalt = 0, desiredt = 0, tstep = 0, runtime = 0, omega = 0;
    mcm::BeadModel model;

    mpcd_ctx *ctx = nullptr;
    mcm_sim_opts opts{};
    bool started = false, has_md = false, done = false;
    int substeps = 0;
    long long steps = 0;
    std::vector<double> pos, vel;
    std::chrono::steady_clock::time_point t_start, t_last_save;
    double globalt_start = 0;

    ~mcm_sim() {
        if (ctx) mpcd_destroy(ctx);
    }

    std::string out(const std::string &file) const { return dir + "/" + file; }

    // World::loadparameter (src/libcilia/global.cpp:21-30) and the object list.
    void load(const std::string &path) {
        cfg_path = path;
        dir = dir_of(path);
        root = mcm::cfg_read_file(path);
        name = root["name"].as_string();
        globalt = root["globalt"].as_double();
        desiredt = root["desiredt"].as_double();
        omega = root["omega"].as_double();  // gravity-driven flow only, which the reference builds with 0
        const char *axes[3] = {"sizex", "sizey", "sizez"};
        for (int d = 0; d < 3; ++d) {
            long long v = root[axes[d]].as_int();
            if (v < 1 || v > 4096) throw CfgError("setting " + std::string(axes[d]) + " is out of range");
            size[d] = (int)v;
        }
        runtime = root["runtime"].as_double();
        tstep = root["tstep"].as_double();
        if (!(tstep > 0.0)) throw CfgError("setting tstep must be positive");
        const Setting &objs = root["md_objects"];
        if (objs.kind != Setting::List && objs.kind != Setting::Group && objs.kind != Setting::Array)
            throw CfgError("setting md_objects is not a list");
        for (int i = 0; i < objs.length(); ++i) {
            const Setting &o = objs[i];
            // entries without a type or a name are skipped (src/simulation.cpp:304-310)
            if (!(o.kind == Setting::Group && o.exists("type") && o["type"].kind == Setting::String &&
                  o.exists("name") && o["name"].kind == Setting::String))
                continue;
            mcm::add_md_object(model, o);
        }
        has_md = model.nbeads() > 0;
    }

    void start(const mcm_sim_opts &o) {
        if (started) throw SimError("simulation already started");
        opts = o;
        if (opts.measureinterval <= 0) opts.measureinterval = 50;
        mpcd_params p{};
        p.size[0] = size[0];
        p.size[1] = size[1];
        p.size[2] = size[2];
        p.rho = kRho;
        p.kT = 1.0;
        p.calpha = -0.6427876;  // src/libcilia/mpcd.h:74-75
        p.salpha = 0.76604444;
        p.precision = opts.precision;
        p.thermostat = 1;
        p.track_ids = 0;
        p.seed = opts.seed ? opts.seed : ((unsigned long long)std::random_device{}() << 32) ^ std::random_device{}();
        opts.seed = p.seed;
        p.tile_variant = -1;
        p.walls = MPCD_WALLS_NONE;
        p.collision = MPCD_SRD;
        check(mpcd_create(&p, &ctx), "mpcd_create");
        const std::string ckpt = out(name + ".ckpt");
        const bool resume = opts.restart && std::ifstream(ckpt).good();
        bool ref_fluid = false;
        if (!resume && opts.restart) {  // the reference's restart files, where a run of it left them
            read_reference_objects();
            if (file_exists(out("mpcd.dat"))) {
                Tokens t(out("mpcd.dat"));
                const double a[3] = {t.num(), t.num(), t.num()};
                check(mpcd_set_shift(ctx, a), "shift from mpcd.dat");
            }
            ref_fluid = file_exists(out("fluid_n.dat"));
            if (ref_fluid) read_reference_fluid();
        }
        if (!resume && !ref_fluid)
            check(mpcd_create_fluid(ctx, (long long)size[0] * size[1] * size[2] * kRho), "create fluid");
        if (has_md) {
            mpcd_md_desc d = model.desc(globalt, kStstep);
            check(mpcd_md_attach(ctx, &d), "mpcd_md_attach");
        }
        if (resume) {
            check(mpcd_checkpoint_load(ctx, ckpt.c_str()), "restart from checkpoint");
            if (has_md) {
                check(mpcd_md_fetch(ctx, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &globalt), "md_fetch");
            } else {
                read_parameter();
            }
        }
        pos.assign(3 * (size_t)model.nbeads(), 0.0);
        vel.assign(4 * (size_t)model.nbeads(), 0.0);
        if (!resume) write_mpcd_debug();
        // Simulation::next_tstep: int((tstep + 1e-8) / ststep) sub-steps (src/simulation.cpp:218)
        substeps = int((tstep + 0.00000001) / kStstep);
        started = true;
        t_start = t_last_save = std::chrono::steady_clock::now();
        globalt_start = globalt;
        if (!opts.quiet)
            std::cout << "Simulation " << name << " (" << size[0] << "x" << size[1] << "x" << size[2] << ", "
                      << model.nbeads() << " beads, seed " << opts.seed << (resume ? ", restarted" : "") << ")"
                      << std::endl
                      << "T=" << globalt << "->" << desiredt << std::endl;
    }

    // MPCD::debug at the start of a run (src/simulation.cpp:30-31,
    // src/libcilia/mpcd.cpp:404-467): one line of mpcd.san over the solvent
    // and the beads: globalt, the dissipation total, particle count, total
    // mass, momentum per mass and sum m v^2 / 3 / count.
    void write_mpcd_debug() {
        const long long n = mpcd_count(ctx);
        std::vector<double> fp(3 * (size_t)n), fv(4 * (size_t)n);
        check(mpcd_download_fluid(ctx, fp.data(), fv.data(), n), "download fluid");
        if (has_md)
            check(mpcd_md_fetch(ctx, nullptr, vel.data(), nullptr, nullptr, nullptr, nullptr, nullptr), "md_fetch");
        double u[3] = {0, 0, 0}, e2 = 0.0;
        long long mass = 0, num = 0;
        auto add = [&](const double *v) {
            const double m = v[3];
            for (int d = 0; d < 3; ++d) u[d] += v[d] * m;
            mass += (long long)m;
            e2 += m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            ++num;
        };
        for (long long i = 0; i < n; ++i) add(&fv[4 * i]);
        for (int b = 0; has_md && b < model.nbeads(); ++b) add(&vel[4 * b]);
        mpcd_diag dg{};
        check(mpcd_diagnostics(ctx, &dg), "mpcd_diagnostics");
        std::ofstream f = append("mpcd.san");
        f << globalt << " " << dg.diss_total << " " << num << " " << mass << " " << u[0] / mass << " " << u[1] / mass
          << " " << u[2] / mass << " " << e2 / 3.0 / num << "\n";
    }

    void need_started() const {
        if (!started) throw SimError("simulation not started (mcm_sim_start)");
    }

    void after_steps(long long n) {
        steps += n;
        if (has_md) {
            check(mpcd_md_fetch(ctx, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, &globalt), "md_fetch");
        } else {
            for (long long s = 0; s < n; ++s)
                for (int k = 0; k < substeps; ++k) globalt += kStstep;  // World::globalt as next_tstep advances it
        }
    }

    void next_tstep(int n) {
        need_started();
        for (int s = 0; s < n; ++s) check(mpcd_step(ctx, tstep, nullptr, nullptr, nullptr), "mpcd_step");
        after_steps(n);
    }

    // Clami::writedata (src/clami.cpp:207-277) and MDObject::writedata
    // (src/libcilia/mdobject.cpp:83-107).
    void writedata() {
        need_started();
        if (!has_md) return;
        check(mpcd_md_fetch(ctx, pos.data(), vel.data(), nullptr, nullptr, nullptr, nullptr, &globalt), "md_fetch");
        for (const mcm::ObjectInfo &o : model.objects) {
            if (o.type == "Clami") write_clami(o);
            else if (o.write_data) write_rods(o);
        }
    }

    std::ofstream append(const std::string &file) {
        std::ofstream f(out(file), std::ios::out | std::ios::app);
        if (!f.good()) throw SimError("cannot write " + out(file));
        return f;
    }

    void write_rods(const mcm::ObjectInfo &o) {
        std::ofstream f = append(o.name + ".xyz");
        f << o.nbeads << "\n" << "rods: " << globalt << "\n";
        for (int b = o.first; b < o.first + o.nbeads; ++b)
            f << "C " << pos[3 * b] << " " << pos[3 * b + 1] << " " << pos[3 * b + 2] << "\n";
    }

    void write_clami(const mcm::ObjectInfo &o) {
        const mpcd_md_sphere &sp = model.sph[o.sphere];
        double xcm[3] = {0, 0, 0}, vcm[3] = {0, 0, 0};
        for (int b = o.first; b < o.first + o.nbeads; ++b)  // sphere, then each cilium rod by rod
            for (int d = 0; d < 3; ++d) {
                xcm[d] += pos[3 * b + d];
                vcm[d] += vel[4 * b + d];
            }
        {
            std::ofstream f = append(o.name + ".xyz");
            const int R = o.cilia.empty() ? 0 : model.cil[o.cilia[0]].rodlength;
            const int even = R % 2;
            f << o.cilia.size() * 3 * (size_t)(R / 2 + even) << "\n" << "rods: " << globalt << "\n";
            for (int ci : o.cilia) {
                const mpcd_md_cilium &c = model.cil[ci];
                for (int rod = 0; rod < 3; ++rod)
                    for (int j = 0; j < c.rodlength; ++j)
                        if ((j + 1) % 2 == even) {
                            const int b = c.first + rod * c.rodlength + j;
                            f << (rod == 0 ? "H " : "O ") << pos[3 * b] << " " << pos[3 * b + 1] << " "
                              << pos[3 * b + 2] << "\n";
                        }
            }
        }
        const int nump = o.nbeads;
        for (int d = 0; d < 3; ++d) {
            xcm[d] /= nump;
            vcm[d] /= nump;
        }
        auto axis = [&](int k, int d) { return pos[3 * (sp.first + k) + d] - pos[3 * sp.first + d]; };
        mpcd_diag dg{};
        check(mpcd_diagnostics(ctx, &dg), "mpcd_diagnostics");
        std::ofstream f = append(o.name + ".san");
        f << globalt << " " << xcm[0] << " " << xcm[1] << " " << xcm[2] << " " << vcm[0] << " " << vcm[1] << " "
          << vcm[2];
        for (int k : {1, 32, 137})
            for (int d = 0; d < 3; ++d) f << " " << axis(k, d);
        f << " " << dg.diss_total / opts.measureinterval / tstep << "\n";
        check(mpcd_reset_diss(ctx), "mpcd_reset_diss");
    }

    void read_parameter() {
        std::ifstream in(out("parameter"));
        std::string key;
        double v;
        while (in >> key >> v)
            if (key == "globalt") {
                globalt = v;
                return;
            }
        throw SimError("restart: no globalt in " + out("parameter"));
    }

    // ---- the reference's restart files (text; read where a run of the
    // reference left them, written at a save when opts.ref_files)

    // fluid_n.dat (Fluid::FluidFromFile / Fluid::load, src/libcilia/fluid.cpp:21-48,121-150)
    void read_reference_fluid() {
        Tokens t(out("fluid_n.dat"));
        const long long n = tagged_int(t, "#numberofparticles");
        const long long want = (long long)size[0] * size[1] * size[2] * kRho;
        if (n > want) throw SimError("fluid_n.dat: Fluidfile to big.");
        if (n < 1) throw SimError("fluid_n.dat: no particles");
        std::vector<double> p(3 * (size_t)n), v(4 * (size_t)n);
        for (double &x : p) x = t.num();
        t.tag("#velocities");
        for (double &x : v) x = t.num();
        check(mpcd_upload_fluid(ctx, p.data(), v.data(), n), "fluid from fluid_n.dat");
    }

    // <name>.dat of each object (MDObject::load, Sphere::load, Cilia::load,
    // ConstantBeatCilia, Clami::load: src/libcilia/mdobject.cpp:110-121,
    // src/sphere.cpp:63-85, src/cilia.cpp:57-66, src/ConstantBeatCilia.h:14-21,
    // src/clami.cpp:304-345), into the bead model before it is attached.
    void read_reference_objects() {
        std::vector<mpcd_md_attachment> att;
        for (mcm::ObjectInfo &o : model.objects) {
            const std::string path = out(o.name + ".dat");
            if (o.type == "Clami") {
                const size_t a0 = att.size();
                if (file_exists(path)) {
                    Tokens t(path);
                    t.tag("#springs");
                    const long long nc = t.integer();
                    if (nc != (long long)o.cilia.size()) throw SimError(path + ": Mismatch between created cilia and dat file.");
                    const int sfirst = model.sph[o.sphere].first;
                    long long ci = t.integer();
                    for (long long i = 0; i < nc; ++i) {
                        if (ci != i) throw SimError(path + ": Mismatch between created cilia and dat file index.");
                        const mpcd_md_cilium &c = model.cil[o.cilia[i]];
                        while (ci == i) {
                            const long long sp = t.integer(), cp = t.integer();
                            const double l0 = t.num();
                            if (sp < 0 || sp >= model.sph[o.sphere].npart || cp < 0 || cp >= 3 * c.rodlength)
                                throw SimError(path + ": spring index out of range");
                            att.push_back(mpcd_md_attachment{sfirst + (int)sp, c.first + (int)cp, l0});
                            std::string w;
                            if (!(t.in >> w)) { ci = -1; break; }
                            ci = std::strtoll(w.c_str(), nullptr, 10);
                        }
                    }
                } else {
                    att.insert(att.end(), model.att.begin() + o.att_first, model.att.begin() + o.att_first + o.att_count);
                }
                o.att_first = (int)a0;
                o.att_count = (int)(att.size() - a0);
                continue;
            }
            if (!file_exists(path)) continue;
            Tokens t(path);
            const bool sphere = o.type == "Sphere";
            const long long rods = tagged_int(t, "#numberofrods"), len = tagged_int(t, "#rodlength");
            if (rods * len != o.nbeads) throw SimError(path + ": object size differs from the configuration");
            for (int b = o.first; b < o.first + o.nbeads; ++b)
                for (int d = 0; d < 3; ++d) model.pos3[3 * (size_t)b + d] = t.num();
            t.tag("#velocities");
            for (int b = o.first; b < o.first + o.nbeads; ++b) {
                for (int d = 0; d < 3; ++d) model.vel4[4 * (size_t)b + d] = t.num();
                t.num();  // the mass column: MDObject::init sets partmass (src/libcilia/mdobject.cpp:152-159)
            }
            if (sphere) {
                // Sphere::load keeps the file's neighbour table (sphere.cpp:72-78)
                // and recalcforce springs bead i to row i's first neilist[i][0]
                // entries (sphere.cpp:111-116): the table built from the
                // configuration must be that one, or the forces would differ
                t.tag("#neilist");
                const mpcd_md_sphere &s = model.sph[o.sphere];
                for (int i = 0; i < o.nbeads; ++i) {
                    long long row[7];
                    for (long long &x : row) x = t.integer();
                    const int b = model.nei_start[s.nei_off + i], e = model.nei_start[s.nei_off + i + 1];
                    bool same = i > 0 ? row[0] == e - b : true;  // the centre's row is not read
                    for (int j = 0; same && j < row[0] && i > 0; ++j) same = row[1 + j] == model.nei[b + j];
                    if (!same)
                        throw SimError(path + ": #neilist row " + std::to_string(i) +
                                       " differs from the neighbour table built from the configuration");
                }
                t.tag("#headbond");
                model.sph[o.sphere].headbond = t.num();
            } else {
                mpcd_md_cilium &c = model.cil[o.cilia[0]];
                size_t off = 0;
                for (int k = 0; k < o.cilia[0]; ++k) off += (size_t)model.cil[k].rodlength;
                t.tag("#oldlength");
                for (int i = 0; i < c.rodlength; ++i) model.oldlength[off + i] = t.num();
                t.tag("#midpnt");
                c.midpnt = t.num();
                t.tag("#tstroke");
                c.Tstroke = t.num();
            }
        }
        model.att = att;
    }

    // The same files from the present state, every number with all its
    // digits (the reference's fluid_n.dat keeps 7: std::scientific).
    void write_reference_files() {
        const long long n = mpcd_count(ctx);
        std::vector<double> fp(3 * (size_t)n), fv(4 * (size_t)n);
        check(mpcd_download_fluid(ctx, fp.data(), fv.data(), n), "download fluid");
        {
            std::ofstream f(out("fluid_n.dat"));
            f << "#numberofparticles " << n << "\n";
            char buf[192];  // four numbers of at most 25 characters each
            for (long long i = 0; i < n; ++i) {
                std::snprintf(buf, sizeof buf, "%.17e %.17e %.17e\n", fp[3 * i], fp[3 * i + 1], fp[3 * i + 2]);
                f << buf;
            }
            f << "\n#velocities\n";
            for (long long i = 0; i < n; ++i) {
                std::snprintf(buf, sizeof buf, "%.17e %.17e %.17e %.17e\n", fv[4 * i], fv[4 * i + 1], fv[4 * i + 2],
                              fv[4 * i + 3]);
                f << buf;
            }
            f << "\n";
            if (!f.good()) throw SimError("cannot write " + out("fluid_n.dat"));
        }
        {
            double a[3];
            check(mpcd_get_shift(ctx, a), "shift");
            std::ofstream f(out("mpcd.dat"));
            f << full(a[0]) << " " << full(a[1]) << " " << full(a[2]) << "\n";
        }
        if (!has_md) return;
        std::vector<double> beat(2 * std::max<size_t>(1, model.cil.size())), nl(std::max<size_t>(1, model.oldlength.size())),
            ol(nl.size());
        check(mpcd_md_fetch(ctx, pos.data(), vel.data(), nullptr, beat.data(), nl.data(), ol.data(), nullptr), "md_fetch");
        for (const mcm::ObjectInfo &o : model.objects) {
            std::ofstream f(out(o.name + ".dat"));
            if (o.type == "Clami") {
                f << "#springs\n" << o.cilia.size() << "\n";
                const int sfirst = model.sph[o.sphere].first;
                for (int k = o.att_first; k < o.att_first + o.att_count; ++k) {
                    const mpcd_md_attachment &a = model.att[k];
                    size_t ci = 0;
                    while (ci + 1 < o.cilia.size() && a.bead1 >= model.cil[o.cilia[ci + 1]].first) ++ci;
                    f << ci << " " << a.bead0 - sfirst << " " << a.bead1 - model.cil[o.cilia[ci]].first << " " << full(a.l0)
                      << "\n";
                }
                continue;
            }
            const bool sphere = o.type == "Sphere";
            const int rods = sphere ? 1 : 3, len = o.nbeads / rods;
            f << "#numberofrods " << rods << "\n#rodlength " << len << "\n";
            for (int r = 0; r < rods; ++r) {
                for (int j = 0; j < len; ++j) {
                    const int b = o.first + r * len + j;
                    f << full(pos[3 * b]) << " " << full(pos[3 * b + 1]) << " " << full(pos[3 * b + 2]) << "\n";
                }
                f << "\n";
            }
            f << "#velocities\n";
            for (int r = 0; r < rods; ++r) {
                for (int j = 0; j < len; ++j) {
                    const int b = o.first + r * len + j;
                    f << full(vel[4 * b]) << " " << full(vel[4 * b + 1]) << " " << full(vel[4 * b + 2]) << " "
                      << full(vel[4 * b + 3]) << "\n";
                }
                f << "\n";
            }
            if (sphere) {
                const mpcd_md_sphere &sp = model.sph[o.sphere];
                f << "#neilist\n";
                for (int i = 0; i < sp.npart; ++i) {
                    const int *row = model.nei_start.data() + sp.nei_off;
                    const int cnt = row[i + 1] - row[i];
                    f << cnt;
                    for (int k = 0; k < 6; ++k) f << " " << (k < cnt ? model.nei[row[i] + k] : 0);
                    f << "\n";
                }
                f << "#headbond\n" << full(sp.headbond) << "\n";
            } else {
                const int ci = o.cilia[0];
                size_t off = 0;
                for (int k = 0; k < ci; ++k) off += (size_t)model.cil[k].rodlength;
                f << "#oldlength\n";
                for (int i = 0; i < model.cil[ci].rodlength; ++i) f << (i ? " " : "") << full(ol[off + i]);
                f << "\n\n#midpnt\n" << full(beat[2 * ci]) << "\n#tstroke\n" << full(beat[2 * ci + 1]);
            }
            if (!f.good()) throw SimError("cannot write " + out(o.name + ".dat"));
        }
    }

    // <name>.ckpt and World::saveparameter's `parameter` file
    // (src/libcilia/global.cpp:52-67), with every digit of globalt.
    void save() {
        need_started();
        check(mpcd_checkpoint_save(ctx, out(name + ".ckpt").c_str()), "checkpoint");
        std::ofstream f(out("parameter"));
        char buf[64];
        std::snprintf(buf, sizeof buf, "%.17g", globalt);
        f << "globalt " << buf << "\n"
          << "desiredt " << desiredt << "\n"
          << "runtime " << runtime << "\n"
          << "sizex " << size[0] << "\nsizey " << size[1] << "\nsizez " << size[2] << "\n"
          << "omega " << omega << "\n";
        if (!f.good()) throw SimError("cannot write " + out("parameter"));
        if (opts.ref_files) write_reference_files();
    }

    double elapsed(std::chrono::steady_clock::time_point since) const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - since).count();
    }

    bool reached_end() const {
        if (opts.max_steps > 0) return steps >= opts.max_steps;
        return !(globalt < desiredt);
    }

    // After a block of measureinterval steps: outputs, progress line and the
    // save/stop logic of src/simulation.cpp:105-141.  Returns true to stop.
    bool end_of_block(double block_seconds) {
        writedata();
        if (!opts.quiet) {
            std::cout << name << " T =" << globalt;
            if (block_seconds > 0) std::cout << " (" << tstep * opts.measureinterval / block_seconds << "fps)";
            std::cout << std::endl;
        }
        if (elapsed(t_start) > runtime) return true;
        if (elapsed(t_last_save) > kSaveEvery) {
            if (!opts.quiet) std::cout << "save stats" << std::endl;
            save();
            t_last_save = std::chrono::steady_clock::now();
        }
        return reached_end();
    }

    // The end of a run: save (Simulation::~Simulation) and write the file's
    // globalt back (src/main.cpp:94-96).
    void finish() {
        save();
        root.set_double("globalt", globalt);
        mcm::cfg_write_file(root, cfg_path);
        if (!opts.quiet) {
            const double wall = elapsed(t_start);
            std::cout << "end simulation: " << name << " " << globalt - globalt_start << " in " << wall << " s";
            if (wall > 0) std::cout << " (" << steps / wall << " steps/s)";
            std::cout << std::endl;
        }
        done = true;
    }

    void run() {
        need_started();
        if (done) throw SimError("simulation already ran to its end");
        while (!reached_end()) {
            const auto t0 = std::chrono::steady_clock::now();
            next_tstep(opts.measureinterval);
            check(mpcd_sync(ctx), "mpcd_sync");
            if (end_of_block(elapsed(t0))) break;
        }
        finish();
    }
};

extern "C" {

const char *mcm_sim_last_error(void) { return g_err.c_str(); }

int mcm_sim_load(const char *cfg_path, mcm_sim **out) {
    if (!cfg_path || !out) {
        g_err = "mcm_sim_load: null argument";
        return MPCD_ERR_ARG;
    }
    *out = nullptr;
    return guarded([&] {
        mcm_sim *s = new mcm_sim;
        try {
            s->load(cfg_path);
        } catch (...) {
            delete s;
            throw;
        }
        *out = s;
    });
}

int mcm_sim_free(mcm_sim *s) {
    delete s;
    return MPCD_OK;
}

int mcm_sim_get_info(const mcm_sim *s, mcm_sim_info *o) {
    if (!s || !o) {
        g_err = "mcm_sim_get_info: null argument";
        return MPCD_ERR_ARG;
    }
    *o = mcm_sim_info{};
    copy_str(o->name, sizeof o->name, s->name);
    for (int d = 0; d < 3; ++d) o->size[d] = s->size[d];
    o->globalt = s->globalt;
    o->desiredt = s->desiredt;
    o->tstep = s->tstep;
    o->runtime = s->runtime;
    o->omega = s->omega;
    o->nobjects = (int)s->model.objects.size();
    o->nbeads = s->model.nbeads();
    o->nspheres = (int)s->model.sph.size();
    o->ncilia = (int)s->model.cil.size();
    o->nattach = (int)s->model.att.size();
    o->steps = s->steps;
    return MPCD_OK;
}

int mcm_sim_get_object(const mcm_sim *s, int i, mcm_sim_object *o) {
    if (!s || !o || i < 0 || i >= (int)s->model.objects.size()) {
        g_err = "mcm_sim_get_object: bad argument";
        return MPCD_ERR_ARG;
    }
    const mcm::ObjectInfo &m = s->model.objects[i];
    *o = mcm_sim_object{};
    copy_str(o->type, sizeof o->type, m.type);
    copy_str(o->name, sizeof o->name, m.name);
    o->write_data = m.write_data ? 1 : 0;
    o->first = m.first;
    o->nbeads = m.nbeads;
    o->sphere = m.sphere;
    o->ncilia = (int)m.cilia.size();
    o->first_cilium = m.cilia.empty() ? -1 : m.cilia[0];
    o->first_attach = m.att_first;
    o->nattach = m.att_count;
    return MPCD_OK;
}

int mcm_sim_md_desc(const mcm_sim *s, mpcd_md_desc *out) {
    if (!s || !out) {
        g_err = "mcm_sim_md_desc: null argument";
        return MPCD_ERR_ARG;
    }
    *out = s->model.desc(s->globalt, kStstep);
    return MPCD_OK;
}

int mcm_sim_start(mcm_sim *s, const mcm_sim_opts *opts) {
    if (!s) {
        g_err = "mcm_sim_start: null simulation";
        return MPCD_ERR_ARG;
    }
    mcm_sim_opts o{};
    if (opts) o = *opts;
    if (o.precision != MPCD_FP64 && o.precision != MPCD_FP32BULK) {
        g_err = "mcm_sim_start: unknown precision";
        return MPCD_ERR_ARG;
    }
    return guarded([&] { s->start(o); });
}

int mcm_sim_next_tstep(mcm_sim *s, int n) {
    if (!s || n < 0) {
        g_err = "mcm_sim_next_tstep: bad argument";
        return MPCD_ERR_ARG;
    }
    return guarded([&] { s->next_tstep(n); });
}

int mcm_sim_writedata(mcm_sim *s) {
    if (!s) {
        g_err = "mcm_sim_writedata: null simulation";
        return MPCD_ERR_ARG;
    }
    return guarded([&] { s->writedata(); });
}

int mcm_sim_run(mcm_sim *s) {
    if (!s) {
        g_err = "mcm_sim_run: null simulation";
        return MPCD_ERR_ARG;
    }
    return guarded([&] { s->run(); });
}

int mcm_sim_save(mcm_sim *s) {
    if (!s) {
        g_err = "mcm_sim_save: null simulation";
        return MPCD_ERR_ARG;
    }
    return guarded([&] { s->save(); });
}

mpcd_ctx *mcm_sim_context(mcm_sim *s) { return s ? s->ctx : nullptr; }

int mcm_sim_run_ensemble(mcm_sim *const *sims, int n) {
    if (!sims || n <= 0) {
        g_err = "mcm_sim_run_ensemble: no simulations";
        return MPCD_ERR_ARG;
    }
    return guarded([&] {
        for (int i = 0; i < n; ++i) {
            if (!sims[i]) throw SimError("mcm_sim_run_ensemble: null simulation");
            sims[i]->need_started();
            if (sims[i]->done) throw SimError("mcm_sim_run_ensemble: " + sims[i]->name + " already ran to its end");
            if (sims[i]->tstep != sims[0]->tstep)
                throw SimError("mcm_sim_run_ensemble: the simulations differ in tstep");
            if (sims[i]->opts.measureinterval != sims[0]->opts.measureinterval)
                throw SimError("mcm_sim_run_ensemble: the simulations differ in measureinterval");
        }
        std::vector<mcm_sim *> active;
        for (int i = 0; i < n; ++i) {
            if (sims[i]->reached_end()) sims[i]->finish();
            else active.push_back(sims[i]);
        }
        const int block = sims[0]->opts.measureinterval;
        while (!active.empty()) {
            std::vector<mpcd_ctx *> boxes;
            for (mcm_sim *s : active) boxes.push_back(s->ctx);
            mpcd_ensemble *e = nullptr;
            check(mpcd_ensemble_create(boxes.data(), (int)boxes.size(), &e), "mpcd_ensemble_create");
            std::vector<mcm_sim *> still;
            try {
                check(mpcd_ensemble_set_graph(e, 1), "mpcd_ensemble_set_graph");
                check(mpcd_ensemble_set_md_wide(e, 1), "mpcd_ensemble_set_md_wide");
                bool changed = false;
                while (!changed) {
                    const auto t0 = std::chrono::steady_clock::now();
                    check(mpcd_ensemble_step(e, active[0]->tstep, block), "mpcd_ensemble_step");
                    check(mpcd_ensemble_sync(e), "mpcd_ensemble_sync");
                    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                    still.clear();
                    for (mcm_sim *s : active) {
                        s->after_steps(block);
                        if (s->end_of_block(secs)) changed = true;
                        else still.push_back(s);
                    }
                }
            } catch (...) {
                mpcd_ensemble_destroy(e);
                throw;
            }
            check(mpcd_ensemble_destroy(e), "mpcd_ensemble_destroy");
            for (mcm_sim *s : active)
                if (std::find(still.begin(), still.end(), s) == still.end()) s->finish();
            active = still;
        }
    });
}

}  // extern "C"
