// Bead-spring MD sub-steps on the GPU: the MD half of one MPCD step.
//
// Every bead's force is a sum over an ordered list of terms, built on the
// host in the order the reference's object loops add them (Sphere,
// Cilia, then Clami's attachments), so the FP64 sums are those of the
// reference's loops.  Per force evaluation the beat's rest lengths are
// refreshed when the reference's 0.01-time-unit schedule says so; then each
// bead's terms are evaluated in parallel (4 lanes per bead, a wave or a
// workgroup for a long list) and added in list order, and the bead kicks and drifts
// (double-buffered positions: terms read only the current buffer); then
// the beat state machine.  Each phase of each evaluation is one launch over
// the GPU, for one swimmer or every swimmer of an ensemble, replayed from a
// captured graph; an ensemble can also run a swimmer per workgroup.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstring>

#include <vector>

namespace mcm {

enum : int {
    kTermConst = 0,     // rest length b
    kTermNewLen = 1,    // rest length newlength[idx] (idx: into the swimmer's rest-length array)
    kTermDiag = 2,      // rest length sqrt(b^2 + b newlength[idx]), b the cilium's bond length
    kTermKindMask = 3,
    kTermAttach = 4,    // attachment spring: no force on coincident points
    kTermAttachB = 8,   // its second bead: force -= (x_partner - x_self) fac
};

// One force term, 16 B: the partner bead, a word holding the kind flags
// (bits 0-3), the spring constant's index in the swimmer's table of
// distinct constants (bits 4-11) and the rest-length index (bits 12-31),
// and the rest length (or the cilium's bond length).  The term lists are
// read at every force evaluation.
struct MdTerm {
    int partner;
    uint32_t meta;
    double b;
};
constexpr int kTermKShift = 4, kTermKBits = 8, kTermIdxShift = 12;
constexpr int kTermMaxK = 1 << kTermKBits;                  // distinct spring constants per swimmer
constexpr int kTermMaxIdx = 1 << (32 - kTermIdxShift);     // rest-length entries per swimmer
__host__ __device__ inline int term_flags(const MdTerm &T) { return (int)(T.meta & 15u); }
__host__ __device__ inline int term_kidx(const MdTerm &T) { return (int)((T.meta >> kTermKShift) & (kTermMaxK - 1)); }
__host__ __device__ inline int term_idx(const MdTerm &T) { return (int)(T.meta >> kTermIdxShift); }

struct MdCilium {
    int first, rodlength, num_anker, constant_beat;
    int len_off;  // offset of the cilium's rest-length arrays
    int pad;
    double bondlength, beatstrength, bending, min_bend, max_bend, m_speed, springconstant;
    double power_stroke_duration, recovery_stroke_duration;
    double midpnt, Tstroke;
};

constexpr int kMdMaxSubsteps = 1024;  // MD sub-steps per MPCD step

// Force evaluations m = 0 .. n at which the beat's rest lengths are refreshed.
struct MdSchedule {
    uint32_t w[kMdMaxSubsteps / 32];
};

// Captured graphs of one step's launches, one per refresh schedule.  The
// schedules of a step length repeat with a short period (steps of 0.025 in
// sub-steps of 0.001: two; 0.1 in 0.003: ten), so each is captured once and
// replayed after.  Past kMax distinct schedules one more graph, which
// launches the refresh at every evaluation (the refresh kernel checks each
// swimmer's own schedule, so the result is the same), serves the rest.
struct SchedGraphs {
    static constexpr int kMax = 16;
    MdSchedule sched[kMax];
    hipGraph_t graph[kMax + 1] = {};  // [kMax]: the catch-all
    hipGraphExec_t exec[kMax + 1] = {};
    int n = 0;

    void clear()
    {
        for (int i = 0; i <= kMax; ++i) {
            if (exec[i])
                (void)hipGraphExecDestroy(exec[i]);
            if (graph[i])
                (void)hipGraphDestroy(graph[i]);
            exec[i] = nullptr;
            graph[i] = nullptr;
        }
        n = 0;
    }

    // The graph of a step whose refreshes are s (evaluations 0 .. nsub):
    // s's own, or, with kMax others held, the catch-all; capture(sched, g,
    // x) captures and instantiates a missing one.
    template <class Capture>
    hipError_t get(const MdSchedule &s, int nsub, Capture capture, hipGraphExec_t &out)
    {
        for (int i = 0; i < n; ++i)
            if (std::memcmp(&sched[i], &s, sizeof s) == 0) {
                out = exec[i];
                return hipSuccess;
            }
        const int slot = n < kMax ? n : kMax;
        if (slot == kMax && exec[kMax]) {
            out = exec[kMax];
            return hipSuccess;
        }
        MdSchedule use = s;
        if (slot == kMax) {
            use = MdSchedule{};
            for (int m = 0; m <= nsub && m < kMdMaxSubsteps; ++m)
                use.w[m >> 5] |= 1u << (m & 31);
        }
        const hipError_t e = capture(use, graph[slot], exec[slot]);
        if (e != hipSuccess)
            return e;
        if (slot < kMax)
            sched[n++] = s;
        out = exec[slot];
        return hipSuccess;
    }
};

// Workgroups of each phase of the wide path, per swimmer: prefix sums over
// the swimmers (nsw + 1 entries each) and the light-bead workgroups of each
// swimmer (4 lanes per bead; its other bead workgroups take, first, the
// beads with the longest term lists, a workgroup each, then the other
// beads with long lists, a wave each).
struct MdWide {
    const int *cil_blk, *bead_blk, *light_blk;
    int cil_total, bead_total;
    const int *ref_blk;  // the rest-length refresh: kMdRefLanes threads per cilium
    int ref_total;
    int nmax;        // most sub-steps of any swimmer
    int nsw;
    int block;       // threads per workgroup of the wide launches
};
// The wide launches' workgroup sizes (an ensemble's autotuner picks one);
// 512 is the default.
constexpr int kMdWideBlock = 512;
constexpr int kMdWideBlocks[3] = {256, 512, 1024};
inline bool md_block_ok(int b) { return b == 256 || b == 512 || b == 1024; }

struct MdBox;

struct MdSystem {
    int nb = 0, nterms = 0, ncil = 0, nlen = 0;
    double *x[2] = {nullptr, nullptr};
    int cur = 0;
    double *v = nullptr;
    double *f = nullptr;
    unsigned char *mobile = nullptr;
    int *tstart = nullptr;
    MdTerm *terms = nullptr;
    double *ktab = nullptr;  // the distinct spring constants of the terms
    MdCilium *cil = nullptr;
    double *newlen = nullptr, *oldlen = nullptr;
    double globalt = 0.0, ststep = 0.001;
    // beads with more terms than a light bead's lanes take, longest list
    // first: the first nbig a workgroup each, the rest a wave each
    int nheavy = 0, nbig = 0;
    int *heavy = nullptr;
    int walls = 0;               // planar walls of the solvent box (bounce-back in the drift)
    double ztop = 0.0, vw[2] = {0.0, 0.0};
    // md_run: the swimmer as a one-swimmer ensemble on the device (MdBox,
    // wide tables), its MD half replayed from a graph per refresh schedule
    MdBox *dbox = nullptr;
    int *dtab = nullptr;
    MdWide wide{};
    SchedGraphs graphs;
    double graph_tstep = -1.0;   // < 0: rebuilt from the host state at the next md_run
};

// A swimmer's device arrays as one kernel argument.
struct MdArgs {
    int nb, nterms, ncil, nheavy;
    double *v, *f;
    const unsigned char *mobile;
    const int *tstart;
    const MdTerm *terms;
    const double *ktab;
    MdCilium *cil;
    double *newlen, *oldlen;
    const int *heavy;
    int nbig;
};

// One swimmer of an ensemble, on the device: its arrays, sub-steps and
// running state (position buffer, time), which the batched MD launch
// advances so that replays of a captured graph continue it.
struct MdBox {
    MdArgs M;
    double *x[2];
    int cur;
    int n;           // MD sub-steps per MPCD step
    double dt;       // sub-step
    double globalt;
    int walls;
    double ztop, vw[2];
    // the wide path: this step's refresh schedule and end time
    MdSchedule sched;
    double g_end;
};


// Sub-steps of one MPCD step of length tstep (simulation.cpp:217).
inline int md_substeps(double ststep, double tstep)
{
    return int((tstep + 0.00000001) / ststep);
}

// The reference's globalt at force evaluation m = 0 .. n: t0, t0, then
// t0 + dt, t0 + 2 dt, ... by repeated addition (simulation.cpp:217-246);
// Cilia::recalcforce refreshes the rest lengths when
// fmod(globalt + ststep / 10, 0.01) < ststep (cilia.cpp:207).  Marks those
// evaluations in sched and returns globalt after the step.  The same code
// runs on the host (md_run) and on the device (ensembles); fmod is exact.
__host__ __device__ inline double md_schedule(double g, double dt, int n, bool has_cilia, MdSchedule &sched)
{
    for (int i = 0; i < kMdMaxSubsteps / 32; ++i)
        sched.w[i] = 0u;
    for (int m = 0; m <= n; ++m) {
        if (m >= 2)
            g += dt;
        if (has_cilia && fmod(g + dt / 10., 0.01) < dt)
            sched.w[m >> 5] |= 1u << (m & 31);
    }
    return g + dt;  // after the last sub-step (simulation.cpp:246)
}

hipError_t md_alloc_upload(MdSystem &s, const double *pos3, const double *vel4, const unsigned char *mobile,
                           const std::vector<int> &tstart, const std::vector<MdTerm> &terms,
                           const std::vector<double> &ktab, const std::vector<MdCilium> &cil,
                           const std::vector<double> &newlen,
                           const std::vector<double> &oldlen, hipStream_t st);
void md_free(MdSystem &s);
// The MD half of one MPCD step (src/simulation.cpp:217-265 after the
// collision): the wide path, replayed from a graph captured at the first
// run with this tstep.  Advances s.globalt as the reference does.
hipError_t md_run(MdSystem &s, double tstep, hipStream_t st);
MdArgs md_args(const MdSystem &s);
// A swimmer's device state for steps of length tstep, from the host's.
MdBox md_box(const MdSystem &s, double tstep);
// The wide path's workgroup tables of nsw swimmers for workgroups of block
// threads (tab goes to the device; md_wide_bind points w at that copy).
void md_wide_tables(const MdBox *boxes, int nsw, std::vector<int> &tab, MdWide &w, int block = kMdWideBlock);
void md_wide_bind(MdWide &w, const int *dtab);
// What one MPCD step does to the host's view of s (position buffer, time),
// for steps the device ran from an ensemble.
void md_advance_host(MdSystem &s, double tstep);
// The MD half of one step of every swimmer of an ensemble, a workgroup each.
void md_launch_multi(MdBox *boxes, int nsw, double tstep, hipStream_t st);
// The same, each phase of each force evaluation as one launch over every
// swimmer's items (the wide path of md_run for all swimmers together).
// only: launch the refresh phase at the evaluations it marks (a single
// swimmer whose schedule the host knows); nullptr: at every evaluation,
// each workgroup checking its swimmer's schedule.
void md_launch_multi_wide(MdBox *boxes, const MdWide &w, double tstep, hipStream_t st,
                          const MdSchedule *only = nullptr);

}  // namespace mcm
