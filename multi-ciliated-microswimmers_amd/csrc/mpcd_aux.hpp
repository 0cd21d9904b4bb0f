// Launchers of the kernels around the step (defined in mpcd_aux.hip).
#pragma once

#include "mpcd_kernels.hpp"
#include "mpcd_md.hpp"

namespace mcm {

// Beads per swimmer of an ensemble (their binning sorts in LDS).
constexpr int kEnsMaxBeads = 16384;

// One box of an ensemble, on the device: its step arguments for both
// buffer parities and its running state, which the ensemble kernels
// advance, so that a captured graph replays whole steps.
struct EnsBox {
    StepArgs tmpl[2];  // cur = 0 / 1; the per-step fields are filled in each step
    int *ovf_cnt2;     // the box's two overflow counters (one per step parity)
    uint64_t step;
    double a[3];
    int cur;
    const MdBox *md;   // the box's swimmer, or null (then tmpl[].nb = 0)
};

void fill_i32(int *p, int v, long long n, hipStream_t s);
void fill_u64(void *p, unsigned long long v, long long n, hipStream_t s);
void launch_upload(const Geom &g, bool f32, bool track, const double *pos3, const double *vel4, long long n,
                   const double a[3], uint32_t *home, int *home_cnt, int *flags, hipStream_t s);
// Applies the step's impulses to the rigid sphere and moves it (one thread).
void launch_sphere_finish(SphereDev *sp, double h, const int L[3], hipStream_t s);
// sum3: kCreateSum3 doubles of scratch (the mean and the blocks' partial sums)
constexpr int kCreateBlocks = 2048;
constexpr int kCreateSum3 = 4 + 3 * kCreateBlocks;
void launch_create_fluid(const double ext[3], double kT, uint64_t seed, long long n, double *pos3, double *vel4,
                         double *sum3, hipStream_t s);
// The same fluid written straight into the home segments under the shift a
// (no staging arrays): the partial sums, the mean and the records; sum3 is
// kCreateSum3 doubles of scratch as above.
void launch_create_records(const Geom &g, bool f32, bool track, const double ext[3], double kT, uint64_t seed,
                           long long n, double *sum3, const double a[3], uint32_t *home, int *home_cnt, int *flags,
                           hipStream_t s);
void launch_fix_overflow(const Geom &g, bool f32, bool track, const uint32_t *ovf, const int *ovf_cnt,
                         int *ovf_cnt_next, const double an[3], uint32_t *home, int *home_cnt, int *flags,
                         hipStream_t s);
void launch_tile_totals(const Geom &g, const int *home_cnt, const int *ob_cnt, int *tot, hipStream_t s);
void launch_download(const Geom &g, bool f32, bool track, const uint32_t *home, const uint32_t *ob, const int *home_cnt,
                     const int *ob_cnt, const long long *prefix, const double a[3], double *pos3, double *vel4,
                     hipStream_t s);
// Checkpoints: the records as stored (tile order), and back into home
// segments under the shift a.
void launch_download_raw(const Geom &g, bool f32, bool track, const uint32_t *home, const uint32_t *ob,
                         const int *home_cnt, const int *ob_cnt, const long long *prefix, uint32_t *words,
                         hipStream_t s);
void launch_upload_raw(const Geom &g, bool f32, bool track, const uint32_t *words, long long n, const double a[3],
                       uint32_t *home, int *home_cnt, int *flags, hipStream_t s);
size_t bead_sort_temp_bytes(int nb_max);
// Beads binned for the next collision; mobile (device, may be null) marks
// the beads that join it (fixed beads, mobile == 0, do not).
int launch_bead_bin(const Geom &g, int nc, const double *bpos3, const double *bvel4, const unsigned char *mobile,
                    int nb, const double a[3], unsigned *keys, unsigned *vals, unsigned *skeys, unsigned *svals,
                    void *temp, size_t temp_bytes, int *b_lc, int *b_idx, double *b_vel4, int *tile_start,
                    hipStream_t s);
void launch_begin_step(Diag *d, bool reset_run, hipStream_t s);
// An ensemble step: launch_ens_begin (nbox workgroups), launch_step_multi
// over every box's tiles, launch_ens_fix, then md_launch_multi.
void launch_ens_begin(EnsBox *boxes, StepArgs *args, int nbox, hipStream_t s);
void launch_ens_fix(EnsBox *boxes, const StepArgs *args, int nbox, bool f32, bool track, hipStream_t s);

}  // namespace mcm
