// Launchers of the kernels around the step (defined in mpcd_aux.hip).
#pragma once

#include "mpcd_kernels.hpp"

namespace mcm {

void fill_i32(int *p, int v, long long n, hipStream_t s);
void fill_u64(void *p, unsigned long long v, long long n, hipStream_t s);
void launch_upload(const Geom &g, bool f32, bool track, const double *pos3, const double *vel4, long long n,
                   const double a[3], uint32_t *home, int *home_cnt, int *flags, hipStream_t s);
// Applies the step's impulses to the rigid sphere and moves it (one thread).
void launch_sphere_finish(SphereDev *sp, double h, const int L[3], hipStream_t s);
void launch_create_fluid(const double ext[3], double kT, uint64_t seed, long long n, double *pos3, double *vel4,
                         double *sum3, hipStream_t s);
void launch_fix_overflow(const Geom &g, bool f32, bool track, const uint32_t *ovf, const int *ovf_cnt,
                         int *ovf_cnt_next, const double an[3], uint32_t *home, int *home_cnt, int *flags,
                         hipStream_t s);
void launch_tile_totals(const Geom &g, const int *home_cnt, const int *ob_cnt, int *tot, hipStream_t s);
void launch_download(const Geom &g, bool f32, bool track, const uint32_t *home, const uint32_t *ob, const int *home_cnt,
                     const int *ob_cnt, const long long *prefix, const double a[3], double *pos3, double *vel4,
                     hipStream_t s);
size_t bead_sort_temp_bytes(int nb_max);
int launch_bead_bin(const Geom &g, int nc, const double *bpos3, const double *bvel4, int nb, const double a[3],
                    unsigned *keys, unsigned *vals, unsigned *skeys, unsigned *svals, void *temp, size_t temp_bytes,
                    int *b_lc, int *b_idx, double *b_vel4, int *tile_start, hipStream_t s);
void launch_begin_step(Diag *d, bool reset_run, hipStream_t s);

}  // namespace mcm
