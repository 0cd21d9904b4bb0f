// The `cilia` executable: `cilia sim.cfg` runs the configuration as the
// reference's src/main.cpp:118-141 does.  Several files run together as one
// ensemble on the GPU.  Options (none in the reference):
//   --fp32            FP32-bulk solvent records (default: FP64, the reference's)
//   --seed N          key of the random streams (default: drawn)
//   --steps N         stop after N steps instead of at desiredt
//   --interval N      steps between outputs (default 50, SIM_MEASUREINTERVAL)
//   --fresh           ignore an existing <name>.ckpt and the reference's restart files
//   --ref-files       saves also write the reference's restart files (fluid_n.dat, ...)
//   --quiet           no progress lines
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mcm_sim.h"

namespace {

int usage(const char *argv0) {
    std::fprintf(stderr,
                 "usage %s [--fp32] [--seed N] [--steps N] [--interval N] [--fresh] [--ref-files] [--quiet] <configfile> "
                 "[<configfile> ...]\n",
                 argv0);
    return 1;
}

int fail(const char *what) {
    std::fprintf(stderr, "%s: %s\n", what, mcm_sim_last_error());
    return 1;
}

bool parse_ll(const char *s, long long &out) {
    char *end = nullptr;
    out = std::strtoll(s, &end, 10);
    return end && *end == '\0' && out >= 0;
}

}  // namespace

int main(int argc, char **argv) {
    mcm_sim_opts opts{};
    opts.precision = MPCD_FP64;
    opts.measureinterval = 50;
    opts.restart = 1;
    std::vector<std::string> files;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        long long v = 0;
        if (a == "--fp32") {
            opts.precision = MPCD_FP32BULK;
        } else if (a == "--fresh") {
            opts.restart = 0;
        } else if (a == "--quiet") {
            opts.quiet = 1;
        } else if (a == "--ref-files") {
            opts.ref_files = 1;
        } else if ((a == "--seed" || a == "--steps" || a == "--interval") && i + 1 < argc) {
            if (!parse_ll(argv[++i], v)) return usage(argv[0]);
            if (a == "--seed") opts.seed = (unsigned long long)v;
            else if (a == "--steps") opts.max_steps = v;
            else if (v < 1 || v > 1000000) return usage(argv[0]);
            else opts.measureinterval = (int)v;
        } else if (!a.empty() && a[0] == '-') {
            return usage(argv[0]);
        } else {
            files.push_back(a);
        }
    }
    if (files.empty()) return usage(argv[0]);

    std::vector<mcm_sim *> sims;
    int rc = 0;
    for (size_t k = 0; k < files.size() && rc == 0; ++k) {
        mcm_sim *s = nullptr;
        if (mcm_sim_load(files[k].c_str(), &s) != MPCD_OK) {
            rc = fail(files[k].c_str());
            break;
        }
        sims.push_back(s);
        mcm_sim_opts o = opts;
        if (o.seed) o.seed += k;  // one stream per box
        if (mcm_sim_start(s, &o) != MPCD_OK) rc = fail(files[k].c_str());
    }
    if (rc == 0) {
        if (sims.size() == 1) {
            if (mcm_sim_run(sims[0]) != MPCD_OK) rc = fail(files[0].c_str());
        } else if (mcm_sim_run_ensemble(sims.data(), (int)sims.size()) != MPCD_OK) {
            rc = fail("ensemble");
        }
    }
    for (mcm_sim *s : sims) mcm_sim_free(s);
    return rc;
}
