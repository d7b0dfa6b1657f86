/*
 * batoid_amd.h -- C ABI of the MI355X (gfx950) trace + Huygens-PSF library.
 *
 * Every entry point takes plain pointers and sizes.  Ray columns, surface,
 * medium and obscuration descriptors, plans and outputs are DEVICE pointers
 * owned by the caller (the Python host keeps them alive); the library never
 * allocates or frees caller memory.  Calls are asynchronous on the HIP stream
 * passed as `stream` (NULL = the legacy default stream); the caller
 * synchronises before reading results.  Each call returns 0 on success, a
 * negative BT_E_* code for an argument error, or a positive hipError_t for a
 * launch failure; bt_last_error() returns a message for the most recent
 * failure on the calling thread.
 *
 * Per-ray failure is NOT an error: it is reported through the `failed` and
 * `vignetted` flag columns with the reference's exact semantics
 * (src/batoid.cpp:236-268, 478-487 of jmeyers314/batoid).
 *
 * The reference boundary replaced here is the pybind11 module
 * `batoid._batoid` (pysrc/batoid.cpp:31-154).  Each function below cites the
 * reference function whose role it takes.
 */
#ifndef BATOID_AMD_H
#define BATOID_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BT_ABI_VERSION 5

/* ---- status codes ------------------------------------------------------ */
#define BT_OK            0
#define BT_E_NULL       -1   /* a required pointer is NULL                  */
#define BT_E_SIZE       -2   /* a size or count is out of range             */
#define BT_E_KIND       -3   /* unknown surface/medium/obscuration kind     */
#define BT_E_PLAN       -4   /* malformed plan                              */
#define BT_E_UNSUPPORTED -5  /* valid request this build does not implement */

/* ---- ray columns --------------------------------------------------------
 * Structure-of-arrays, FP64 except the two flag columns (1 byte each, 0/1).
 * Replaces CPPRayVector(x_ptr, ..., failed_ptr, size)
 * (pysrc/rayVector.cpp:31-62; include/rayVector.h:8-39).
 * Velocities are in units of c (|v| = 1/n); t is optical path in metres.  */
typedef struct bt_rays {
    double* x; double* y; double* z;
    double* vx; double* vy; double* vz;
    double* t;
    double* wavelength;
    double* flux;
    uint8_t* vignetted;
    uint8_t* failed;
    int64_t n;
} bt_rays;

/* ---- surfaces -----------------------------------------------------------
 * One record per surface; a Sum record is immediately followed by the
 * records of its `nchild` children in order, each a leaf record or a Sum
 * with its own children (nested at most 3 levels below a Sum), and its n1
 * is the number of records below it (so a child Sum's subtree can be
 * stepped over).  Derived constants are
 * filled by the host exactly as the reference constructors compute them.
 *   PLANE        plane.cpp:13-30
 *   SPHERE       p = {R, R*R, 1/R, 1/R/R}                 sphere.cpp:9-72
 *   PARABOLOID   p = {R, 1/R, 1/2/R}                      paraboloid.cpp:9-71
 *   QUADRIC      p = {R, k, R*R, 1/R/R, k+1, 1/(k+1), R/(k+1),
 *                     R*R/(k+1)/(k+1), (k+1)/R/R}         quadric.cpp:9-77
 *   ASPHERE      p as QUADRIC; n1 = ncoef; a0 = coefs (index 0 <-> r^2),
 *                a1 = dz/dr coefs 2(j+1)*coefs[j]         asphere.cpp:10-85
 *   TILTED       p = {tanx, tany}                         tilted.cpp:14-37
 *   BICUBIC      p = {x0, y0, dx, dy}; n1 = nx, n2 = ny; flags = use_nan;
 *                a0..a3 = z, dzdx, dzdy, d2zdxdy row-major [ny][nx]
 *                                                         table.cpp:58-158
 *   POLYNOMIAL   p = {xorigin, yorigin}; n1 = xsize, n2 = ysize;
 *                a0 = coefs[ysize][xsize], a1 = gradx, a2 = grady
 *                ([ysize-1][xsize-1])                     polynomialSurface.cpp:35-66
 *                a3 = optional int32 trim table (or NULL): for a0, a1, a2
 *                in turn, the last row holding a coefficient other than
 *                +0.0 (-1 if none), then each row's length up to its last
 *                such coefficient (1 + n2, n2, n2 entries).  Evaluation
 *                skips the +0.0 runs it marks, with bit-identical results.
 *   SUM          nchild children follow; n1 = records below   sum.cpp:23-56
 */
enum bt_surface_kind {
    BT_PLANE = 0, BT_SPHERE = 1, BT_PARABOLOID = 2, BT_QUADRIC = 3,
    BT_ASPHERE = 4, BT_TILTED = 5, BT_BICUBIC = 6, BT_POLYNOMIAL = 7,
    BT_SUM = 8
};

typedef struct bt_surface {
    int32_t kind;
    int32_t nchild;
    int32_t n1, n2;
    int32_t flags;
    int32_t pad_;
    double p[12];
    const double* a0;
    const double* a1;
    const double* a2;
    const double* a3;
} bt_surface;

/* ---- media --------------------------------------------------------------
 *   CONST      p = {n}                                    medium.cpp:30-32
 *   TABLE      n = size; a0 = wavelengths, a1 = indices   medium.cpp:89-105
 *   SELLMEIER  p = {B1, B2, B3, C1, C2, C3}               medium.cpp:145-149
 *   SUMITA     p = {A0, ..., A5}                          medium.cpp:186-191
 *   AIR        p = {P*7.50061683, T-273.15, W*7.50061683,
 *                   P, T, W}  (kPa, K, kPa inputs kept)   medium.cpp:229-245
 */
enum bt_medium_kind {
    BT_MED_CONST = 0, BT_MED_TABLE = 1, BT_MED_SELLMEIER = 2,
    BT_MED_SUMITA = 3, BT_MED_AIR = 4
};

typedef struct bt_medium {
    int32_t kind;
    int32_t n;
    double p[6];
    const double* a0;
    const double* a1;
} bt_medium;

/* ---- obscurations -------------------------------------------------------
 * An obscuration is a postfix program of `len` consecutive records; the
 * value left on the stack is `contains(x, y)`.
 *   CIRCLE     p = {radius, x0, y0}                       obscuration.cpp:32-34
 *   ANNULUS    p = {inner, outer, x0, y0}                 obscuration.cpp:80-83
 *   RECTANGLE  p = {width, height, x0, y0, cos th, sin th, th}
 *                                                         obscuration.cpp:117-121
 *   RAY        p = {width, x0, y0, cos th, sin th, th}    obscuration.cpp:155-159
 *   POLYGON    n = size; a0 = xs, a1 = ys                 obscuration.cpp:201-225
 *   NEGATION   pops 1                                     obscuration.cpp:320-322
 *   UNION      pops n                                     obscuration.cpp:364-370
 *   INTERSECTION pops n                                   obscuration.cpp:415-421
 */
enum bt_obsc_op {
    BT_OB_CIRCLE = 0, BT_OB_ANNULUS = 1, BT_OB_RECTANGLE = 2, BT_OB_RAY = 3,
    BT_OB_POLYGON = 4, BT_OB_NEGATION = 5, BT_OB_UNION = 6,
    BT_OB_INTERSECTION = 7
};
#define BT_OBSC_MAX_STACK 32

typedef struct bt_obsc {
    int32_t op;
    int32_t n;
    double p[8];
    const double* a0;
    const double* a1;
} bt_obsc;

/* ---- plans --------------------------------------------------------------
 * A plan is the flattened Optic.trace of a CompoundOptic
 * (batoid/optic.py:391-424, 1013-1049): one step per Interface, each
 * carrying the coordinate transform from the previous frame
 * (coordTransform.py:16-20; drot row-major as numpy's ravel, trace.py:76),
 * the interaction, the surface, the outgoing medium and the obscuration.  */
enum bt_action { BT_ACT_INTERSECT = 0, BT_ACT_REFLECT = 1, BT_ACT_REFRACT = 2 };

/* How a step seeds the Newton intersection of a Newton surface (Asphere,
 * Bicubic, Zernike/Polynomial, Sum):
 *   REFERENCE  the reference's start and exactly niter iterations
 *              (surface.cpp:22-49, asphere.cpp:75-85, sum.cpp:47-56):
 *              bit-identical to the reference;
 *   CONIC      an analytic start, then Newton on the full surface until the
 *              sag gap is <= 1e-14 m (at most niter iterations).  The start:
 *              an Asphere's quadric root; for a Sum led by an Asphere, that
 *              Asphere's quadric root moved by one Newton update on the
 *              Asphere; for any other Sum, the reference's own start (its
 *              first child's root, sum.cpp:53); the ray origin for a
 *              polynomial or Bicubic surface;
 *   MLP        the base conic's root (a Sum's first leaf's; the ray origin
 *              without one) plus the learned correction of the step's
 *              network, then the same iteration.
 * A ray that does not converge from a fast start within niter iterations is
 * re-intersected on the REFERENCE path, which decides its flags.  Where the
 * fast start is the reference's own start (CONIC on an Asphere, and on a
 * Sum not led by an Asphere: the same iterates), a ray the reference fails
 * is therefore failed here too.  From a different start (CONIC on a Sum led
 * by an Asphere, an MLP prediction) a fast start can converge on a ray
 * whose reference iteration would not have converged in niter steps: that
 * ray is intersected where the reference marks it failed.  Positions agree
 * with the reference to <= 1e-12 m where both intersect.  Starts other than
 * REFERENCE are honoured only by bt_trace_plan with BT_PLAN_FAST_START set;
 * MLP starts only with BT_PLAN_MLP_START also set, and an MLP step with a
 * null network runs as CONIC. */
enum bt_start { BT_START_REFERENCE = 0, BT_START_CONIC = 1, BT_START_MLP = 2 };

typedef struct bt_step {
    int32_t action;
    int32_t surface;    /* index of the first record in the surface table   */
    int32_t medium;     /* index in the medium table (REFRACT), else -1     */
    int32_t obsc;       /* index of the first obscuration record, or -1     */
    int32_t obsc_len;   /* number of obscuration records                    */
    int32_t start;      /* bt_start                                         */
    double dr[3];
    double drot[9];
    const void* mlp;    /* packed network (bt_mlp_pack) for BT_START_MLP    */
} bt_step;

/* bt_plan.flags */
#define BT_PLAN_NO_REFRACT 1   /* no step refracts: wavelength is not read */
#define BT_PLAN_FAST_START 2   /* some step uses a CONIC or MLP start      */
#define BT_PLAN_MLP_START  4   /* some step uses an MLP start; without this
                                  flag MLP steps run as CONIC             */
#define BT_PLAN_ANALYTIC   8   /* every surface record is a Plane, Sphere,
                                  Paraboloid, Quadric, Asphere or Tilted:
                                  the trace runs a kernel without the
                                  Bicubic, polynomial and Sum code (fewer
                                  registers); under this flag any other kind
                                  evaluates to NaN and fails              */
#define BT_PLAN_FLAT_SUMS 32   /* (without the two flags above) no Sum has
                                  a Sum child: the trace runs a kernel
                                  without the nested-Sum evaluation (fewer
                                  registers); under this flag a Sum child
                                  of a Sum evaluates to NaN and fails      */
#define BT_PLAN_SUM_APB   16   /* (without BT_PLAN_ANALYTIC) every surface
                                  record is one of the kinds above or a Sum
                                  of exactly (Asphere, Polynomial, Bicubic)
                                  children in that order -- a perturbed
                                  conic: the trace runs a kernel with those
                                  three children unrolled; under this flag
                                  a Bicubic or Polynomial outside such a
                                  Sum, or a Sum of another layout,
                                  evaluates to NaN and fails              */

typedef struct bt_plan {
    int32_t nstep;
    int32_t flags;
    int32_t nmedia;               /* records in the medium table             */
    int32_t pad_;
    const bt_step* steps;         /* device */
    const bt_surface* surfaces;   /* device */
    const bt_medium* media;       /* device */
    const bt_obsc* obscs;         /* device */
} bt_plan;

/* ---- library ----------------------------------------------------------- */
int bt_abi_version(void);
const char* bt_last_error(void);
/* sizeof of bt_rays, bt_surface, bt_medium, bt_obsc, bt_step, bt_plan,
 * bt_psf_segment, bt_raygen, bt_coating, in that order, written to sizes[0..n-1]; lets a binding
 * check its layout. */
int bt_struct_sizes(int64_t* sizes, int n);
/* Field layout of struct number `which` (the order of bt_struct_sizes):
 * bt_struct_field_count gives its number of fields (BT_E_SIZE for an
 * unknown struct), bt_struct_field_offset the byte offset of the field
 * named `field` as spelled in this header (-1 if it has no such field).
 * With bt_struct_sizes a binding can check each of its fields by name, so
 * two same-size fields declared in the wrong order are caught. */
int bt_struct_field_count(int which);
int64_t bt_struct_field_offset(int which, const char* field);
/* Device properties the host uses for launch sizing. */
int bt_device_info(int* cu_count, int* arch_major, int* arch_minor);

/* ---- coordinate transforms ---------------------------------------------
 * Replace _batoid.applyForwardTransform / applyReverseTransform
 * (src/batoid.cpp:71-148, pysrc/batoid.cpp:53-54) and the *Arrays forms
 * (src/batoid.cpp:8-38, pysrc/batoid.cpp:61-102).  dr[3], drot[9] are HOST
 * arrays (passed by value to the kernel).                                  */
int bt_apply_forward_transform(const double* dr, const double* drot,
                               bt_rays rays, void* stream);
int bt_apply_reverse_transform(const double* dr, const double* drot,
                               bt_rays rays, void* stream);
int bt_apply_forward_transform_arrays(const double* dr, const double* drot,
                                      double* x, double* y, double* z,
                                      int64_t n, void* stream);
int bt_apply_reverse_transform_arrays(const double* dr, const double* drot,
                                      double* x, double* y, double* z,
                                      int64_t n, void* stream);
/* Replaces _batoid.finishParallel (src/batoid.cpp:40-69,
 * pysrc/batoid.cpp:103-121): rays on the stop surface are moved back onto
 * the plane through `dr` perpendicular to `vv`.                            */
int bt_finish_parallel(const double* dr, const double* drot, const double* vv,
                       double* x, double* y, double* z, int64_t n,
                       void* stream);

/* ---- ray generation -----------------------------------------------------
 * Replaces the body of RayVector.asGrid / asPolar / fromStop
 * (batoid/rayVector.py:219-444, 459-643, 860-982) for a parallel beam: stop
 * coordinates, the stop surface's sag, the stop -> ray frame transform
 * (src/batoid.cpp:8-22), finishParallel (src/batoid.cpp:40-69) and the
 * constant columns (rayVector.py:831-841), all in one pass that writes the
 * 11 columns of `out` (out.n rays).
 *   POINTS         stop (x, y) = (px[i], py[i])                 fromStop
 *   GRID           ray i = (iy, ix) = (i / nx, i % nx) of the
 *                  fftshift(fftfreq) lattice: u = (ix - nx/2) fx,
 *                  w = (iy - ny/2) fy; (x, y) = (lx.(u, w), ly.(u, w))
 *                                                 rayVector.py:418-433
 *   GRID_RANDOM    u = lo[0] + span[0] U_i, w = lo[1] + span[1] U_(n+i)
 *                                                 rayVector.py:410-416
 *   POLAR          ring k holds rays [ring_off[k], ring_off[k+1]); ray j of
 *                  the ring at r = ring_rho[k], th = j ring_dth[k]
 *                                                 utils.py:512-531
 *   POLAR_RANDOM   r = sqrt(lo[0] + span[0] U_i), th = lo[1] + span[1] U_(n+i)
 *                                                 rayVector.py:625-630
 * U_k is the k-th double (k from 0) a PCG64 (XSL-RR 128/64) generator with
 * state pcg_state / increment pcg_inc yields, (next64 >> 11) * 2^-53: the
 * stream numpy.random.Generator(PCG64).uniform draws, so a seeded
 * default_rng gives the same rays as the reference.  POLAR* take
 * (r cos th, r sin th) with the device's FP64 sin/cos.
 * Every pointer is a device pointer; `flux` may be NULL (all rays flux0). */
enum bt_gen_pattern {
    BT_GEN_POINTS = 0, BT_GEN_GRID = 1, BT_GEN_GRID_RANDOM = 2,
    BT_GEN_POLAR = 3, BT_GEN_POLAR_RANDOM = 4
};

typedef struct bt_raygen {
    int32_t pattern;
    int32_t nring;              /* POLAR: rings in the ring tables          */
    int64_t nx, ny;             /* GRID: lattice shape, nx * ny == out.n    */
    double fx, fy;              /* GRID: fftfreq spacing 1 / (n d)          */
    double lx[2], ly[2];        /* GRID*: lattice edge vectors              */
    double lo[2], span[2];      /* *_RANDOM: uniform(lo, lo + span)          */
    uint64_t pcg_state[2];      /* *_RANDOM: PCG64 state {hi, lo}           */
    uint64_t pcg_inc[2];        /*           increment {hi, lo}             */
    const int64_t* ring_off;    /* POLAR: nring + 1 ray offsets             */
    const double* ring_rho;     /* POLAR: nring radii                        */
    const double* ring_dth;     /* POLAR: nring azimuth steps                */
    const double* px;           /* POINTS: stop coordinates                  */
    const double* py;
    const double* flux;         /* NULL or out.n per-ray fluxes              */
    const bt_surface* stop;     /* stop surface record                       */
    double stop_dr[3], stop_drot[9];    /* stop frame -> ray frame           */
    double origin[3], rot[9];   /* finishParallel plane (rayVector.py:822-829) */
    double vv[3];               /* ray velocity (rayVector.py:820-821)       */
    double wavelength, flux0;
} bt_raygen;

int bt_generate_rays(const bt_raygen* gen, bt_rays out, void* stream);

/* ---- coatings ----------------------------------------------------------
 * SimpleCoating (src/coating.cpp:10-39), the reference's only coating:
 * constant reflected and transmitted flux fractions.  Host struct, passed by
 * pointer and copied into the launch.                                      */
typedef struct bt_coating {
    double reflectivity;
    double transmissivity;
} bt_coating;

/* ---- per-surface trace --------------------------------------------------
 * Replace _batoid.intersect / reflect / refract (src/batoid.cpp:175-506,
 * pysrc/batoid.cpp:55-57) with the caller's niter.  `surface` is a device
 * pointer to the surface's first record; `m2` a device pointer to the
 * outgoing medium (m1 is not needed: the reference takes n1 from |v|,
 * src/batoid.cpp:454-458).  `coating` (host pointer, NULL = none)
 * multiplies the flux of every ray that reaches the surface in this call by
 * its transmissivity (intersect, refract -- total internal reflection
 * included) or reflectivity (reflect), as src/batoid.cpp:251-263, 362-371,
 * 497-499 do.  `obsc`/`obsc_len` optionally fold the Interface's
 * obscuration into the same pass (optic.py:419-422); pass NULL/0 for the
 * plain reference call.                                                    */
int bt_intersect(const bt_surface* surface, const double* dr,
                 const double* drot, bt_rays rays, int niter,
                 const bt_coating* coating, const bt_obsc* obsc, int obsc_len,
                 void* stream);
int bt_reflect(const bt_surface* surface, const double* dr,
               const double* drot, bt_rays rays, int niter,
               const bt_coating* coating, const bt_obsc* obsc, int obsc_len,
               void* stream);
int bt_refract(const bt_surface* surface, const double* dr,
               const double* drot, const bt_medium* m2, bt_rays rays,
               int niter, const bt_coating* coating, const bt_obsc* obsc,
               int obsc_len, void* stream);
/* bt_intersect (BT_ACT_INTERSECT), bt_reflect (BT_ACT_REFLECT) or
 * bt_refract (BT_ACT_REFRACT, m2 used) for ray sets that traceSplit's
 * filter (optic.py:1256-1266) has been applied to without removing the
 * rays it drops: rays with vignetted != 0 or flux < min_flux (NaN
 * included) are left unchanged; every other ray is processed exactly as
 * by the plain call.  `kinds` and `counts` as for bt_rsplit_live, with one
 * count spread over counts[0..63].  No reference counterpart (see
 * bt_rsplit_live).                                                        */
int bt_interact_live(int action, const bt_surface* surface, const double* dr,
                     const double* drot, const bt_medium* m2, bt_rays rays, int niter,
                     const bt_coating* coating, const bt_obsc* obsc, int obsc_len,
                     double min_flux, int kinds, int64_t* counts, void* stream);
/* Replaces _batoid.rSplit (src/batoid.cpp:509-674, pysrc/batoid.cpp:58):
 * each ray that reaches the surface is split into the refracted ray
 * (written back to `rays`, flux x transmissivity) and the reflected ray
 * (written to `split`, flux x reflectivity).  Rays already failed on entry
 * are copied to `split` unchanged (the reference's rvSplit starts as a copy
 * of rv, batoid/trace.py:125); rays that miss the surface are
 * marked failed and vignetted in both.  `split` must hold rays.n rays.
 * `obsc`/`obsc_len` optionally fold the Interface's obscuration into the
 * same pass for both parts (optic.py:491-494, as traceSplit applies it
 * after rSplit); pass NULL/0 for the plain reference call.                */
int bt_rsplit(const bt_surface* surface, const double* dr, const double* drot,
              const bt_medium* m2, const bt_coating* coating, const bt_obsc* obsc,
              int obsc_len, bt_rays rays, bt_rays split, int niter, void* stream);
/* bt_rsplit for ray sets that traceSplit's filter (optic.py:1256-1266) has
 * already been applied to without removing the rays it drops: rays with
 * vignetted != 0 or flux < min_flux (NaN included) are left unchanged and
 * only their split part is marked vignetted (its other columns are not
 * written); every other ray is split exactly as by bt_rsplit, except that
 * a split (reflected) ray the same filter drops is written only as its
 * flux, vignetted and failed.  Filtering the two outputs afterwards gives
 * the rays bt_rsplit on the filtered set and the same filter would give.
 * A `split` with every column NULL and n = 0 asks for no split part: only
 * the refracted part is written (traceSplit's use when it already knows
 * that the filter drops every reflected ray).  `kinds`: BT_PLAN_ANALYTIC when the
 * surface record is a Plane, Sphere, Paraboloid, Quadric, Asphere or Tilted
 * (a kernel without the other kinds' code; under it any other kind fails
 * its rays), else 0.  `counts` (a device int64[2 * 64], or NULL) is
 * incremented by the numbers of refracted (entries 0..63) and split
 * (entries 64..127) rays that pass the filter, each number spread over its
 * 64 entries: their sum is the count.  The
 * reference has no such call: its traceSplit copies out the kept rays
 * after every Interface.                                                  */
int bt_rsplit_live(const bt_surface* surface, const double* dr, const double* drot,
                   const bt_medium* m2, const bt_coating* coating, const bt_obsc* obsc,
                   int obsc_len, bt_rays rays, bt_rays split, double min_flux, int niter,
                   int kinds, int64_t* counts, void* stream);

/* ---- ray selection ------------------------------------------------------
 * The ray filter of CompoundOptic.traceSplit (optic.py:1256-1266): keep the
 * rays with vignetted == 0 and flux >= min_flux, in input order.  Two calls:
 * bt_select_count writes the number kept to *count (a device int64) and the
 * per-block offsets to the workspace; after the caller has read the count
 * and sized `out`, bt_select_scatter (same `in`, `min_flux` and workspace)
 * writes the kept rays.  Workspace: bt_select_workspace_bytes(in.n).      */
int64_t bt_select_workspace_bytes(int64_t n);
int bt_select_count(bt_rays in, double min_flux, int64_t* count, void* workspace,
                    int64_t workspace_bytes, void* stream);
int bt_select_scatter(bt_rays in, double min_flux, bt_rays out, const void* workspace,
                      int64_t workspace_bytes, void* stream);

/* Replaces _batoid.obscure (src/batoid.cpp:151-172, pysrc/batoid.cpp:59). */
int bt_obscure(const bt_obsc* obsc, int obsc_len, bt_rays rays, void* stream);

/* Whole-optic fused trace: every step of `plan` applied to each ray while it
 * sits in registers; reads `in` once and writes `out` once (in == out is
 * the in-place Optic.trace).  Same per-ray results as calling the per-step
 * functions in order.  No single reference function: replaces the loop of
 * CompoundOptic.trace (optic.py:1013-1049) over trace.py:47-116.          */
int bt_trace_plan(bt_plan plan, bt_rays in, bt_rays out, int niter,
                  void* stream);

/* ---- surface / medium / obscuration queries -----------------------------
 * Replace CPPSurface.sag / .normal (pysrc/surface.cpp:11-21),
 * CPPMedium.getN (pysrc/medium.cpp), CPPObsc*.contains
 * (pysrc/obscuration.cpp).  All arrays are device arrays of length n;
 * `normal_out` is [3][n] (nx block, ny block, nz block).                   */
int bt_surface_sag(const bt_surface* surface, const double* x,
                   const double* y, double* out, int64_t n, void* stream);
int bt_surface_normal(const bt_surface* surface, const double* x,
                      const double* y, double* normal_out, int64_t n,
                      void* stream);
int bt_medium_get_n(const bt_medium* medium, const double* wavelength,
                    double* out, int64_t n, void* stream);
int bt_obsc_contains(const bt_obsc* obsc, int obsc_len, const double* x,
                     const double* y, uint8_t* out, int64_t n, void* stream);

/* The square roots every kernel of the trace uses: out[0:n] = sqrt(x) and
 * out[n:2n] = 1/sqrt(x) as the trace computes them (x, out device arrays;
 * out holds 2n doubles).  They must equal IEEE-754 double sqrt and division
 * bit for bit, as the reference's CPU arithmetic does; this entry point
 * exists to test that and has no counterpart in the reference.            */
int bt_check_sqrt(const double* x, double* out, int64_t n, void* stream);

/* ---- RayVector field operations -----------------------------------------
 * Replace CPPRayVector.positionAtTime / propagateInPlace / phase /
 * amplitude / sumAmplitude (src/rayVector.cpp:26-222,
 * pysrc/rayVector.cpp:63-98).  `amp_out` is interleaved complex128 [n][2].
 * bt_sum_amplitude writes one complex128 (2 doubles) to the DEVICE pointer
 * `out`; `workspace` must hold bt_sum_amplitude_workspace_bytes(n) bytes.  */
int bt_position_at_time(bt_rays rays, double t, double* xout, double* yout,
                        double* zout, void* stream);
int bt_propagate_in_place(bt_rays rays, double t, void* stream);
int bt_phase(bt_rays rays, double x, double y, double z, double t,
             double* out, void* stream);
int bt_amplitude(bt_rays rays, double x, double y, double z, double t,
                 double* amp_out, void* stream);
int64_t bt_sum_amplitude_workspace_bytes(int64_t n);
int bt_sum_amplitude(bt_rays rays, double x, double y, double z, double t,
                     int ignore_vignetted, double* out, void* workspace,
                     void* stream);

/* ---- learned Newton start ------------------------------------------------
 * No counterpart in the reference: the networks that seed BT_START_MLP
 * steps, 6 -> 64 -> 64 -> 1 with ReLU, trained in FP32 on intersections
 * the tracer produces.  A network reads, in its step's local frame, the
 * ray's point on the surface's base conic (the ray origin where the surface
 * has no conic base) and its direction, and predicts (dt* - dt_conic): the
 * exact intersection time minus the base root.  Parameters are one FP32
 * vector of bt_mlp_param_count() floats: W1[64][6], b1[64], W2[64][64],
 * b2[64], w3[64], b3.  `in_norm` is in_mean[6], in_inv_std[6] (14 floats
 * with out_mean, out_std for bt_mlp_pack).  Features are [n][6] FP32;
 * targets are (dt* - dt_conic - out_mean) / out_std.                        */
int bt_mlp_param_count(void);
int64_t bt_mlp_packed_bytes(void);
int64_t bt_mlp_grad_workspace_bytes(int64_t batch);
/* Mean squared error over `batch` samples (rows index[b] of features /
 * targets, or rows 0..batch-1 when index is NULL; every index must be a
 * valid row) and its gradient. */
int bt_mlp_loss_grad(const float* params, const float* features, const float* targets,
                     const int32_t* index, int64_t batch, const float* in_norm,
                     float* grad, float* loss, void* workspace, int64_t workspace_bytes,
                     void* stream);
int bt_mlp_adam(float* params, const float* grad, float* m, float* v, int64_t nparam,
                float lr, float beta1, float beta2, float eps, int step, void* stream);
/* FP32 inference, output in normalised units. */
int bt_mlp_forward(const float* params, const float* in_norm, const float* features,
                   int64_t n, float* out, void* stream);
/* Pack FP32 parameters + norm[14] into the fp16 MFMA tables used in the
 * trace (bt_mlp_packed_bytes() bytes at `packed`). */
int bt_mlp_pack(const float* params, const float* norm, void* packed, void* stream);
/* Packed (MFMA) inference, output de-normalised (metres). */
int bt_mlp_forward_packed(const void* packed, const float* features, int64_t n,
                          float* out, void* stream);
/* Training samples from the tracer: for rays about to meet `surface` (frame
 * change dr, drot), the network features, the base-conic root, the
 * reference intersection (niter Newton iterations) and status bits (1:
 * conic root found, 2: intersection converged). */
int bt_newton_samples(const bt_surface* surface, const double* dr, const double* drot,
                      bt_rays rays, int niter, float* features, double* dt_conic,
                      double* dt_exact, uint8_t* status, void* stream);

/* ---- Huygens PSF ---------------------------------------------------------
 * The reference evaluates the image of huygensPSF (analysis.py:181-305) as a
 * Python loop of sumAmplitude calls over every pixel (analysis.py:302-303,
 * rayVector.cpp:169-222).  Pixel (row j, column i) sits at
 *     p = p0 + (i - nx/2) a1 + (j - ny/2) a2,  z = 0,
 * with a1, a2 the lattice's primitive vectors (lattice.py:6-14 and the
 * transposition at analysis.py:296-297).  `amp_out` receives complex128
 * amplitudes [ny][nx][2]; intensity is |amp|^2.
 *
 * bt_huygens_direct: FP64 direct sum, one exp(i phi) per (pixel, ray): the
 *   reference's arithmetic on the GPU (the oracle path for the GEMM).
 * bt_huygens_gemm: the sum factorised as a complex GEMM
 *   A[j][i] = sum_r c_r U[i][r] V[r][j], c_r = flux e^{i phi_r(p0)},
 *   U = e^{i m alpha_r}, V = e^{i n beta_r}, with every phase reduced mod 2pi
 *   in FP64 before the exponentials; the GEMM runs on MFMA with
 *   `precision` = BT_PSF_FP32 (v_mfma_f32 *_f32 operands) or BT_PSF_BF16X3
 *   (split hi+lo bf16 operands, three products, FP32 accumulation).
 *   Both take a device `workspace` of at least
 *   bt_huygens_workspace_bytes(n, nx, ny) bytes for partial sums.           */
enum bt_psf_precision { BT_PSF_FP32 = 0, BT_PSF_BF16X3 = 1 };

int64_t bt_huygens_workspace_bytes(int64_t n, int nx, int ny);
int bt_huygens_direct(bt_rays rays, const double* p0, const double* a1,
                      const double* a2, int nx, int ny, double t,
                      int ignore_vignetted, double* amp_out, void* workspace,
                      int64_t workspace_bytes, void* stream);
int bt_huygens_gemm(bt_rays rays, const double* p0, const double* a1,
                    const double* a2, int nx, int ny, double t,
                    int ignore_vignetted, int precision, double* amp_out,
                    void* workspace, int64_t workspace_bytes, void* stream);

/* Allocation-free bt_huygens_gemm -- no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  Same
 * arguments, kernels and results (bit for bit) as bt_huygens_gemm, but the
 * scratch comes from a fixed device array inside the library
 * (bt_huygens_noalloc_scratch_bytes() = 128 MiB; each GPU the library runs
 * on holds its own copy until the process ends) instead of a caller
 * workspace.  A request whose scratch (bt_huygens_workspace_bytes(n, nx, ny))
 * exceeds it returns BT_E_SIZE before anything is launched.  Every call on
 * one GPU uses that GPU's array: calls that overlap on one GPU (two host
 * threads, or two streams whose work may run concurrently) race and must
 * be serialised by the caller.  Makes no hipMalloc, hipFree or memset call;
 * the first call on each device looks up the array's device address.    */
int64_t bt_huygens_noalloc_scratch_bytes(void);
int bt_huygens_gemm_noalloc(bt_rays rays, const double* p0, const double* a1,
                            const double* a2, int nx, int ny, double t,
                            int ignore_vignetted, int precision, double* amp_out,
                            void* stream);

/* ---- batched Huygens PSFs ---------------------------------------------------
 * One nx x ny PSF per segment of a ray batch, all in one call (the nearest
 * reference is a Python loop of huygensPSF calls, analysis.py:181-305, over
 * (field, wavelength) pairs; concatenation as in rayVector.py:1351-1365).
 * Segment s covers rays [begin, end) (segments disjoint, sorted by begin);
 * its lattice is centred on the position of ray `chief` (chief < 0: the
 * mean position of its unvignetted rays, analysis.py:289-295) at the time t
 * of ray `time_ray` (analysis.py:301), with primitive vectors a1, a2.
 * The segment table lives in device memory, so one call can be captured in a
 * HIP graph and replayed on new ray data: no host synchronisation, no
 * allocation, no memset.  Rays are summed in work items of `chunk` rays (a
 * multiple of 8) by `nwg` persistent workgroups (0: 512) in a fixed
 * assignment; items of one segment are added in FP64 in item order.
 * `amp_out` receives [nseg][ny][nx][2] complex128.                         */
typedef struct bt_psf_segment {
    int64_t begin, end;
    int64_t chief;
    int64_t time_ray;
    double a1[2], a2[2];
} bt_psf_segment;

int64_t bt_huygens_batch_workspace_bytes(int64_t n, int nseg, int nx, int ny, int chunk);
int bt_huygens_batch(bt_rays rays, const bt_psf_segment* segments, int nseg, int nx,
                     int ny, int ignore_vignetted, int precision, int chunk, int nwg,
                     double* amp_out, void* workspace, int64_t workspace_bytes,
                     void* stream);

/* ---- Zernike decomposition of a wavefront --------------------------------
 * The O(N) part of batoid.zernike / zernikeGQ (batoid/analysis.py:479-641;
 * the reference evaluates galsim's zernikeBasis and calls numpy's lstsq or
 * dot on the host).  Over the rays with mask[i] == 0 (mask NULL: all rays):
 *   w_i = (t0 - t[i]) / wavelength            (analysis.py:375, 639)
 *   w_i *= flux[i]                             if flux != NULL (quadrature)
 *   B_j(i) = sum_m coef[j][m] u^(d-q) v^q      u = x/r_outer, v = y/r_outer,
 *                                              m = d(d+1)/2 + q, d <= nmax,
 * for j = 0 .. jmax-1 (Noll j+1), and the Gram matrix of [B_0 .. B_jmax-1, w]
 * is written packed by rows (upper triangle, (jmax+1)(jmax+2)/2 doubles):
 * the normal matrix, the projections sum B_j w and sum w^2.  coef is a
 * device [jmax][(nmax+1)(nmax+2)/2] table of monomial coefficients;
 * jmax <= 78, nmax <= 11.  Summation order is fixed (run-to-run
 * deterministic).  Workspace: bt_zernike_workspace_bytes(jmax) bytes.    */
int64_t bt_zernike_workspace_bytes(int jmax);
int bt_zernike_gram(const double* x, const double* y, const double* t, const double* flux,
                    const uint8_t* mask, int64_t n, double t0, double wavelength,
                    double r_outer, const double* coef, int jmax, int nmax, double* gram,
                    void* workspace, int64_t workspace_bytes, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* BATOID_AMD_H */
