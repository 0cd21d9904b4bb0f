// The MPCD / Fluid facade (include/gpu_mpcd.hpp) driven the way the
// reference's Simulation::next_tstep drives class MPCD (src/simulation.cpp:
// 217-265) with a host-side MD object: a bead chain integrated as
// MDObject::md_loop_start / md_loop / md_loop_stop do (src/libcilia/
// mdobject.cpp:36-72; forces are accelerations, Matrix.h:425-428), its
// beads handed to the collision by add_impuls (MDObject::init and
// md_loop_stop).  The same steps are taken through the C ABI directly
// (mpcd_stage_beads, mpcd_step, mpcd_fetch_beads) on a second context: bead
// states, the solvent and the dissipated energy must agree bit for bit.
// Then: the walls' bounce of MPCD::move on hand-worked cases, the shift kept
// in mpcd.dat across contexts, Fluid::save / Fluid::FluidFromFile, and
// MPCD::debug's line in mpcd.san.  Runs in the current directory; prints
// "facade ok" and exits 0 on success.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "gpu_mpcd.hpp"

namespace {

int failures = 0;

void expect(bool ok, const std::string &what)
{
    if (!ok) {
        std::printf("[facade] FAILED: %s\n", what.c_str());
        ++failures;
    }
}

// A bead chain: positions, velocities (mass 5 in column 3), forces.
struct Chain {
    int nb;
    std::vector<double> x, v, f;
    explicit Chain(int n) : nb(n), x(3 * n), v(4 * n), f(3 * n, 0.0)
    {
        for (int i = 0; i < n; ++i) {
            x[3 * i + 0] = 4.0 + 0.45 * i;
            x[3 * i + 1] = 6.0 + 0.1 * std::sin(0.7 * i);
            x[3 * i + 2] = 5.0 + 0.1 * std::cos(0.3 * i);
            v[4 * i + 0] = 0.05 * std::sin(1.3 * i);
            v[4 * i + 1] = 0.05 * std::cos(0.9 * i);
            v[4 * i + 2] = 0.02;
            v[4 * i + 3] = 5.0;
        }
    }
    // MDObject::spring, one-sided, both directions: bonds of rest 0.5, k 100
    void forces()
    {
        std::fill(f.begin(), f.end(), 0.0);
        for (int i = 0; i + 1 < nb; ++i) {
            for (int s = 0; s < 2; ++s) {
                const int a = s ? i + 1 : i, b = s ? i : i + 1;
                double d2 = 0.0;
                for (int k = 0; k < 3; ++k) {
                    const double d = x[3 * a + k] - x[3 * b + k];
                    d2 += d * d;
                }
                const double fac = -2.0 * 100.0 * (1.0 - 0.5 / std::sqrt(d2));
                for (int k = 0; k < 3; ++k)
                    f[3 * a + k] += (x[3 * a + k] - x[3 * b + k]) * fac;
            }
        }
    }
    void kick(double c)
    {
        for (int i = 0; i < nb; ++i)
            for (int k = 0; k < 3; ++k)
                v[4 * i + k] += f[3 * i + k] * c;
    }
};

void drift_plain(Chain &c, double ts)
{
    for (int i = 0; i < c.nb; ++i)
        for (int k = 0; k < 3; ++k)
            c.x[3 * i + k] += c.v[4 * i + k] * ts;
}

bool same(const std::vector<double> &a, const std::vector<double> &b)
{
    return a.size() == b.size() && std::memcmp(a.data(), b.data(), sizeof(double) * a.size()) == 0;
}

}  // namespace

int main()
{
    const int L = 16, nb = 24, steps = 6;
    const double tstep = 0.05, ststep = 0.001;
    const int n = int((tstep + 0.00000001) / ststep);  // simulation.cpp:218
    std::remove("mpcd.dat");
    std::remove("mpcd.san");
    std::remove("fluid_n.dat");
    mcm::MPCDOptions opt;
    opt.seed = 17;
    opt.tstep = tstep;
    double shift_kept[3];
    long double diss_facade = 0.0L;
    std::vector<double> sol_p, sol_v;
    {
        mcm::MPCD mpcd(L, L, L, opt);
        mcm::Fluid *fluid = mcm::Fluid::FluidFromFile("fluid_n.dat", mpcd);  // no file: Fluid::create
        expect(fluid->count() == 10 * L * L * L, "fluid size");

        mpcd_params p{};
        p.size[0] = p.size[1] = p.size[2] = L;
        p.rho = 10;
        p.kT = 1.0;
        p.calpha = -0.6427876;
        p.salpha = 0.76604444;
        p.precision = MPCD_FP64;
        p.thermostat = 1;
        p.track_ids = 1;
        p.seed = 17;
        p.tile_variant = -1;
        mpcd_ctx *ctx = nullptr;
        expect(mpcd_create(&p, &ctx) == MPCD_OK, "mpcd_create");
        expect(mpcd_create_fluid(ctx, 10LL * L * L * L) == MPCD_OK, "mpcd_create_fluid");

        Chain A(nb), B(nb);
        for (int i = 0; i < nb; ++i)  // MDObject::init
            mpcd.add_impuls(&A.x[3 * i], &A.v[4 * i]);
        std::vector<double> out(4 * nb);
        for (int s = 0; s < steps; ++s) {
            // through the facade, in Simulation::next_tstep's order
            mpcd.next_tstep();
            A.forces();
            A.kick(ststep / 2.0);
            for (int i = 0; i < nb; ++i)
                mpcd.move(&A.x[3 * i], &A.v[4 * i], ststep);
            A.forces();
            for (int k = 1; k < n; ++k) {
                A.kick(ststep);
                for (int i = 0; i < nb; ++i)
                    mpcd.move(&A.x[3 * i], &A.v[4 * i], ststep);
                A.forces();
            }
            mpcd.shift_boxes();
            fluid->move();
            A.kick(ststep / 2.0);
            for (int i = 0; i < nb; ++i)
                mpcd.add_impuls(&A.x[3 * i], &A.v[4 * i]);

            // the same through the C ABI
            expect(mpcd_stage_beads(ctx, B.x.data(), B.v.data(), nb) == MPCD_OK, "stage");
            expect(mpcd_step(ctx, tstep, nullptr, nullptr, nullptr) == MPCD_OK, "step");
            expect(mpcd_fetch_beads(ctx, out.data(), nb) == MPCD_OK, "fetch");
            for (int i = 0; i < nb; ++i)
                for (int k = 0; k < 3; ++k)
                    B.v[4 * i + k] = out[4 * i + k];
            B.forces();
            B.kick(ststep / 2.0);
            drift_plain(B, ststep);
            B.forces();
            for (int k = 1; k < n; ++k) {
                B.kick(ststep);
                drift_plain(B, ststep);
                B.forces();
            }
            B.kick(ststep / 2.0);

            expect(same(A.x, B.x) && same(A.v, B.v), "beads at step " + std::to_string(s));
            double a2[3];
            mpcd_get_shift(ctx, a2);
            expect(std::memcmp(a2, mpcd.shift(), sizeof a2) == 0, "shift at step " + std::to_string(s));
        }
        const long long nf = mpcd_count(ctx);
        std::vector<double> p1(3 * nf), v1(4 * nf), p2(3 * nf), v2(4 * nf);
        expect(mpcd_download_fluid(mpcd.context(), p1.data(), v1.data(), nf) == MPCD_OK, "download facade");
        expect(mpcd_download_fluid(ctx, p2.data(), v2.data(), nf) == MPCD_OK, "download direct");
        expect(same(p1, p2) && same(v1, v2), "solvent after the steps");
        mpcd_diag d{};
        mpcd_diagnostics(ctx, &d);
        expect(std::fabs((double)mpcd.diss_energy_total - d.diss_total) <= 1e-12 * std::fabs(d.diss_total),
               "diss_energy_total");
        expect(mpcd.diss_energy_total != 0.0L, "the collisions dissipate");
        diss_facade = mpcd.diss_energy_total;
        mpcd.diss_energy_total = 0.0;  // Clami::writedata's reset (clami.cpp:281-282)
        mpcd.add_impuls(&A.x[0], &A.v[0]);
        mpcd.next_tstep();
        mpcd_diag d1{};
        mpcd_diagnostics(mpcd.context(), &d1);
        expect(std::fabs((double)mpcd.diss_energy_total - d1.diss_last) <= 1e-9 * std::fabs(d1.diss_last) + 1e-300,
               "after a reset, diss_energy_total is the last step's");
        mpcd.shift_boxes();
        mpcd.debug(0.5);

        fluid->save();
        std::vector<double> p3(3 * nf), v3(4 * nf);
        expect(mpcd_download_fluid(mpcd.context(), p3.data(), v3.data(), nf) == MPCD_OK, "download");
        sol_p = p3;
        sol_v = v3;
        for (int k = 0; k < 3; ++k)
            shift_kept[k] = mpcd.shift()[k];
        delete fluid;
        mpcd_destroy(ctx);
    }  // ~MPCD writes mpcd.dat

    {
        // a new MPCD finds the shift in mpcd.dat; FluidFromFile loads the saved solvent
        mcm::MPCD mpcd(L, L, L, opt);
        expect(std::memcmp(mpcd.shift(), shift_kept, sizeof shift_kept) == 0, "shift read back from mpcd.dat");
        mcm::Fluid *fluid = mcm::Fluid::FluidFromFile("fluid_n.dat", mpcd);
        const long long nf = mpcd_count(mpcd.context());
        std::vector<double> p(3 * nf), v(4 * nf);
        expect(mpcd_download_fluid(mpcd.context(), p.data(), v.data(), nf) == MPCD_OK, "download loaded");
        expect(same(p, sol_p) && same(v, sol_v), "Fluid::save / FluidFromFile round trip");
        fluid->check();
        delete fluid;
    }

    {
        // MPCD::move at the walls (mpcd.cpp:346-377), worked by hand
        mcm::MPCDOptions w = opt;
        w.walls = MPCD_WALLS_NOSLIP;
        w.persist_shift = false;
        w.shift_file = "";
        mcm::MPCD mpcd(8, 8, 8, w);
        double x[3] = {1.0, 1.0, 6.95}, v[4] = {0.5, 0.25, 1.0, 5.0};
        mpcd.move(x, v, 0.1);  // z 7.05 > 7: back along the path for t2 = 0.05
        expect(std::fabs(x[0] - 1.0) < 1e-12 && std::fabs(x[1] - 1.0) < 1e-12 && std::fabs(x[2] - 6.95) < 1e-12,
               "top wall position");
        expect(v[0] == -0.5 && v[1] == -0.25 && v[2] == -1.0, "top wall velocity");
        double y[3] = {1.0, 1.0, 0.05}, u[4] = {0.2, 0.3, -1.0, 5.0};
        mpcd.move(y, u, 0.1);  // z -0.05 < 0: every component reversed
        expect(std::fabs(y[0] - 1.0) < 1e-12 && std::fabs(y[1] - 1.0) < 1e-12 && std::fabs(y[2] - 0.05) < 1e-12,
               "bottom wall position");
        expect(u[0] == -0.2 && u[1] == -0.3 && u[2] == 1.0, "bottom wall velocity");
    }

    std::ifstream san("mpcd.san");
    std::string line;
    int lines = 0, fields = 0;
    while (std::getline(san, line)) {
        ++lines;
        double t;
        fields = 0;
        for (std::istringstream ls(line); ls >> t;)
            ++fields;
    }
    expect(lines == 1 && fields == 8, "mpcd.san line");
    if (failures) {
        std::printf("[facade] %d failures\n", failures);
        return 1;
    }
    std::printf("facade ok: %d steps bit-identical to the C ABI, mpcd.dat, fluid file, walls, debug\n", steps);
    return 0;
}
