/*
 * oracle.h -- CPU restatement of jmeyers314/batoid's trace hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * library (batoid_amd/libbatoid_amd.so) and the CPU baseline timed beside
 * it; only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it.  The product never calls it.
 *
 * It consumes the same POD descriptors as include/batoid_amd.h, but every
 * pointer (ray columns, tables, plans) is a HOST pointer.  Loops over rays
 * are OpenMP-parallel with `nthreads` threads, as the reference's
 * `#pragma omp parallel for num_threads(max_threads)` (src/batoid.cpp:221).
 */
#ifndef BT_ORACLE_H
#define BT_ORACLE_H

#include "../include/batoid_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

double or_sag(const bt_surface* s, double x, double y);
void or_normal(const bt_surface* s, double x, double y,
               double* nx, double* ny, double* nz);
int or_time_to_intersect(const bt_surface* s, double x, double y, double z,
                         double vx, double vy, double vz, double* dt,
                         int niter);
double or_get_n(const bt_medium* m, double wavelength);
int or_contains(const bt_obsc* prog, int len, double x, double y);

void or_sag_n(const bt_surface* s, const double* x, const double* y,
              double* out, int64_t n, int nthreads);
void or_normal_n(const bt_surface* s, const double* x, const double* y,
                 double* out, int64_t n, int nthreads);
void or_get_n_n(const bt_medium* m, const double* w, double* out, int64_t n,
                int nthreads);
void or_contains_n(const bt_obsc* prog, int len, const double* x,
                   const double* y, uint8_t* out, int64_t n, int nthreads);

void or_apply_forward_transform(const double* dr, const double* drot,
                                bt_rays rays, int nthreads);
void or_apply_reverse_transform(const double* dr, const double* drot,
                                bt_rays rays, int nthreads);
void or_apply_forward_transform_arrays(const double* dr, const double* drot,
                                       double* x, double* y, double* z,
                                       int64_t n, int nthreads);
void or_apply_reverse_transform_arrays(const double* dr, const double* drot,
                                       double* x, double* y, double* z,
                                       int64_t n, int nthreads);
void or_finish_parallel(const double* dr, const double* drot, const double* vv,
                        double* x, double* y, double* z, int64_t n);

void or_intersect(const bt_surface* s, const double* dr, const double* drot,
                  bt_rays rays, int niter, int nthreads);
void or_reflect(const bt_surface* s, const double* dr, const double* drot,
                bt_rays rays, int niter, int nthreads);
void or_refract(const bt_surface* s, const double* dr, const double* drot,
                const bt_medium* m2, bt_rays rays, int niter, int nthreads);
void or_intersect_coated(const bt_surface* s, const double* dr, const double* drot,
                         double transmissivity, bt_rays rv, int niter, int nthreads);
void or_reflect_coated(const bt_surface* s, const double* dr, const double* drot,
                       double reflectivity, bt_rays rv, int niter, int nthreads);
void or_refract_coated(const bt_surface* s, const double* dr, const double* drot,
                       const bt_medium* m2, double transmissivity, bt_rays rv, int niter,
                       int nthreads);
void or_rsplit(const bt_surface* s, const double* dr, const double* drot,
               const bt_medium* m2, double reflectivity, double transmissivity,
               bt_rays rv, bt_rays rvs, int niter, int nthreads);
void or_obscure(const bt_obsc* prog, int len, bt_rays rays, int nthreads);
/* Optic.trace as the reference's Python drives it: for each step, the
 * interaction call and then (if any) the obscure call, each a full pass. */
void or_trace_plan(const bt_plan* plan, bt_rays rays, int niter, int nthreads);

void or_position_at_time(bt_rays rays, double t, double* xo, double* yo,
                         double* zo, int nthreads);
void or_propagate_in_place(bt_rays rays, double t, int nthreads);
void or_phase(bt_rays rays, double x, double y, double z, double t,
              double* out, int nthreads);
void or_amplitude(bt_rays rays, double x, double y, double z, double t,
                  double* out, int nthreads);
void or_sum_amplitude(bt_rays rays, double x, double y, double z, double t,
                      int ignore_vignetted, double* out, int nthreads);
/* The huygensPSF pixel loop: one sumAmplitude per pixel (FP64). */
void or_huygens_direct(bt_rays rays, const double* p0, const double* a1,
                       const double* a2, int nx, int ny, double t,
                       int ignore_vignetted, double* amp_out, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
