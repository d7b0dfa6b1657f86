// bt_trace.hip -- ray-surface trace kernels and their C-ABI entry points.
//
// Data layout: the RayVector is 11 structure-of-arrays columns in HBM
// (9 x FP64 + 2 x u8 flags); lane i of a wave handles ray i, so every column
// access is a fully coalesced 512 B (FP64) or 64 B (flag) wave transaction.
// Descriptors (surfaces, media, obscuration programs, plan steps) are read
// through wave-uniform addresses and land in scalar registers.
//
// Two ways to trace:
//   * per-surface kernels (bt_intersect/bt_reflect/bt_refract), one HBM read
//     + write of the touched columns per Interface, with the Interface's
//     obscuration folded into the same pass;
//   * the fused plan kernel (bt_trace_plan), which applies every Interface of
//     an optic while the ray sits in registers: one read and one write per
//     ray for the whole optic.
#include <hip/hip_runtime.h>
#include <string>
#include <cstdio>
#include <cstddef>
#include <cstring>

#include "bt_device.h"
#include "bt_common.h"

namespace bt {

thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg)
{
    set_error(msg);
    return code;
}

int check_launch(const char* what)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error(std::string(what) + ": " + hipGetErrorString(e));
        return (int)e;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------
__device__ __forceinline__ int64_t ray_index()
{
    return (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
}

template <bool kForward>
__global__ __launch_bounds__(kBlock) void k_transform_rays(Xform xf, bt_rays rv)
{                                                    // batoid.cpp:71-148
    const int64_t i = ray_index();
    if (i >= rv.n) return;
    const double* dr = xf.dr;
    const double* m = xf.drot;
    const double x0 = rv.x[i], y0 = rv.y[i], z0 = rv.z[i];
    const double vx0 = rv.vx[i], vy0 = rv.vy[i], vz0 = rv.vz[i];
    if (kForward) {
        double dx = x0-dr[0];
        double dy = y0-dr[1];
        double dz = z0-dr[2];
        rv.x[i] = dx*m[0] + dy*m[3] + dz*m[6];
        rv.y[i] = dx*m[1] + dy*m[4] + dz*m[7];
        rv.z[i] = dx*m[2] + dy*m[5] + dz*m[8];
        rv.vx[i] = vx0*m[0] + vy0*m[3] + vz0*m[6];
        rv.vy[i] = vx0*m[1] + vy0*m[4] + vz0*m[7];
        rv.vz[i] = vx0*m[2] + vy0*m[5] + vz0*m[8];
    } else {
        rv.x[i] = x0*m[0] + y0*m[1] + z0*m[2] + dr[0];
        rv.y[i] = x0*m[3] + y0*m[4] + z0*m[5] + dr[1];
        rv.z[i] = x0*m[6] + y0*m[7] + z0*m[8] + dr[2];
        rv.vx[i] = vx0*m[0] + vy0*m[1] + vz0*m[2];
        rv.vy[i] = vx0*m[3] + vy0*m[4] + vz0*m[5];
        rv.vz[i] = vx0*m[6] + vy0*m[7] + vz0*m[8];
    }
}

template <bool kForward>
__global__ __launch_bounds__(kBlock) void k_transform_arrays(
    Xform xf, double* __restrict__ x, double* __restrict__ y,
    double* __restrict__ z, int64_t n)
{                                                    // batoid.cpp:8-38
    const int64_t i = ray_index();
    if (i >= n) return;
    const double* dr = xf.dr;
    const double* m = xf.drot;
    if (kForward) {
        double dx = x[i]-dr[0];
        double dy = y[i]-dr[1];
        double dz = z[i]-dr[2];
        x[i] = dx*m[0] + dy*m[3] + dz*m[6];
        y[i] = dx*m[1] + dy*m[4] + dz*m[7];
        z[i] = dx*m[2] + dy*m[5] + dz*m[8];
    } else {
        const double x0 = x[i], y0 = y[i], z0 = z[i];
        x[i] = x0*m[0] + y0*m[1] + z0*m[2] + dr[0];
        y[i] = x0*m[3] + y0*m[4] + z0*m[5] + dr[1];
        z[i] = x0*m[6] + y0*m[7] + z0*m[8] + dr[2];
    }
}

struct FinishArgs {
    Xform xf;
    double vl[3];   // -vv rotated into the stop frame (batoid.cpp:44-46)
};

__global__ __launch_bounds__(kBlock) void k_finish_parallel(
    FinishArgs a, double* __restrict__ x, double* __restrict__ y,
    double* __restrict__ z, int64_t n)
{                                                    // batoid.cpp:40-69
    const int64_t i = ray_index();
    if (i >= n) return;
    const double* dr = a.xf.dr;
    const double* m = a.xf.drot;
    double dx = x[i]-dr[0];
    double dy = y[i]-dr[1];
    double dz = z[i]-dr[2];
    double xx = dx*m[0] + dy*m[3] + dz*m[6];
    double yy = dx*m[1] + dy*m[4] + dz*m[7];
    double zz = dx*m[2] + dy*m[5] + dz*m[8];
    double dt = -zz/a.vl[2];
    xx += dt*a.vl[0];
    yy += dt*a.vl[1];
    zz += dt*a.vl[2];
    x[i] = xx*m[0] + yy*m[1] + zz*m[2] + dr[0];
    y[i] = xx*m[3] + yy*m[4] + zz*m[5] + dr[1];
    z[i] = xx*m[6] + yy*m[7] + zz*m[8] + dr[2];
}

// One Interface over all rays, in place.  coat_on / coat: the coating's
// flux factor for rays that reach the surface (batoid.cpp:251-263, 362-371,
// 497-499); a SimpleCoating's factor does not depend on the incidence.
// live / min_flux / counts: as k_rsplit's (traceSplit's carried sets): a
// ray the filter already dropped is skipped, and the kept rays of the
// output are counted into kStepCountSlots partial sums.
constexpr int kStepCountSlots = 64;

template <int K>
__device__ __forceinline__ bool step_ray(
    int64_t i, int action, const Xform& xf, const bt_surface* __restrict__ surf,
    const bt_medium* __restrict__ m2, const bt_obsc* __restrict__ obsc, int obsc_len,
    const bt_rays& rv, int niter, int coat_on, double coat, int live, double min_flux)
{
    Ray r;
    r.fail = rv.failed[i] != 0;
    r.vig = rv.vignetted[i] != 0;
    double f = 0.0;
    if (live) {
        f = rv.flux[i];
        if (r.vig || !(f >= min_flux)) return false;
    }
    const bool fail0 = r.fail, vig0 = r.vig;
    r.x = rv.x[i]; r.y = rv.y[i]; r.z = rv.z[i];
    r.vx = rv.vx[i]; r.vy = rv.vy[i]; r.vz = rv.vz[i];
    r.t = rv.t[i];
    double n2 = 0.0;
    if (action == BT_ACT_REFRACT && !r.fail) {
        r.w = rv.wavelength[i];
        n2 = medium_n(m2, r.w);
    }
    const bool moved = apply_interaction<K>(action, xf.dr, xf.drot, surf, n2, r, niter);
    if (moved) {
        rv.x[i] = r.x; rv.y[i] = r.y; rv.z[i] = r.z;
        rv.vx[i] = r.vx; rv.vy[i] = r.vy; rv.vz[i] = r.vz;
        rv.t[i] = r.t;
        if (coat_on) {
            if (!live) f = rv.flux[i];
            f *= coat;
            rv.flux[i] = f;
        }
    }
    if (obsc_len > 0 && !r.vig)                     // batoid.cpp:169-171
        r.vig = obsc_contains(obsc, obsc_len, r.x, r.y);
    if (r.vig != vig0) rv.vignetted[i] = r.vig;
    if (r.fail != fail0) rv.failed[i] = r.fail;
    return live && !r.vig && f >= min_flux;
}

template <int K>
__global__ __launch_bounds__(kBlock) void k_step(
    int action, Xform xf, const bt_surface* __restrict__ surf,
    const bt_medium* __restrict__ m2, const bt_obsc* __restrict__ obsc,
    int obsc_len, bt_rays rv, int niter, int coat_on, double coat, int live,
    double min_flux, unsigned long long* __restrict__ counts)
{
    const int64_t i = ray_index();
    bool keep = false;
    if (i < rv.n)
        keep = step_ray<K>(i, action, xf, surf, m2, obsc, obsc_len, rv, niter, coat_on, coat,
                           live, min_flux);
    if (counts) {
        const uint64_t m = __ballot(keep);
        if ((threadIdx.x & 63) == 0 && m)
            atomicAdd(counts + (int)(blockIdx.x % kStepCountSlots),
                      (unsigned long long)__popcll(m));
    }
}

__global__ __launch_bounds__(kBlock) void k_obscure(
    const bt_obsc* __restrict__ obsc, int obsc_len, bt_rays rv)
{                                                    // batoid.cpp:151-172
    const int64_t i = ray_index();
    if (i >= rv.n) return;
    if (rv.vignetted[i]) return;
    if (obsc_contains(obsc, obsc_len, rv.x[i], rv.y[i])) rv.vignetted[i] = 1;
}

// Every step of a plan applied while the ray stays in registers.  A ray's
// wavelength never changes, so the index of the plan's (first) dispersive
// medium is evaluated once per ray up front.
//
// Mode 0: reference starts only.  Mode 1: CONIC starts honoured (an MLP step
// runs as CONIC).  Mode 2: MLP starts too; MLP inference is a wave-wide MFMA
// product, so in this variant every lane of a wave stays in the step loop
// (lanes past the end work on a copy of the last ray and store nothing).
//
// The network's input at an MLP step: the ray's point on the surface's base
// conic and its direction, in the step's local frame (the frame change is
// apply_interaction's).
template <int K>
__device__ __forceinline__ void mlp_features(const bt_step* __restrict__ st,
                                             const bt_surface* surf, const Ray& r,
                                             const MlpPacked* P, float (&f)[kMlpIn])
{
    const double* dr = st->dr;
    const double* m = st->drot;
    double dx = r.x-dr[0], dy = r.y-dr[1], dz = r.z-dr[2];
    double x = dx*m[0] + dy*m[3] + dz*m[6];
    double y = dx*m[1] + dy*m[4] + dz*m[7];
    double z = dx*m[2] + dy*m[5] + dz*m[8];
    double vx = r.vx*m[0] + r.vy*m[3] + r.vz*m[6];
    double vy = r.vx*m[1] + r.vy*m[4] + r.vz*m[7];
    double vz = r.vx*m[2] + r.vy*m[5] + r.vz*m[8];
    double dtc = 0.0;
    if (r.fail || !conic_start<K>(surf, x, y, z, vx, vy, vz, dtc)) {
        // keep this lane's share of the wave's evaluation finite
        for (int k = 0; k < kMlpIn; k++) f[k] = P->norm[k];
        return;
    }
    f[0] = (float)(x + vx*dtc);
    f[1] = (float)(y + vy*dtc);
    f[2] = (float)(z + vz*dtc);
    f[3] = (float)vx;
    f[4] = (float)vy;
    f[5] = (float)vz;
}

// Waves per SIMD: the analytic-kinds kernels of modes 0 and 1 are held to
// five (93 VGPRs, no spills), so that a few more registers in a later change
// cannot silently drop them to four; the all-kinds kernel
// of mode 0 is held to four (some registers spill to scratch; the extra
// wave hides more FP64 latency: 8 % faster on the TMA, 9-12 % on the
// Newton-heavy optics), mode 2 to three (left alone the compiler would
// settle for two); the all-kinds kernel of mode 1 keeps the compiler's
// three, where four measured slower on the TMA.
// Networks a plan kernel keeps in LDS (45 KB per workgroup).
constexpr int kMlpLds = 4;

template <int Mode, int K>
constexpr int plan_waves_per_simd()
{
    if (K == kKindsAnalytic && Mode < 2) return 5;
    if (K == kKindsSumAPB) return Mode == 2 ? 3 : 4;
    return Mode == 0 ? 4 : (Mode == 2 ? 3 : 1);     // kKindsAll, kKindsFlatSums
}

template <int Mode, int K>
__global__ __launch_bounds__(kBlock)
__attribute__((amdgpu_waves_per_eu(plan_waves_per_simd<Mode, K>())))
void k_plan(
    bt_plan plan, bt_rays in, bt_rays out, int niter, int need_w)
{
    constexpr bool Fast = Mode > 0;
    constexpr bool Mlp = Mode > 1;
    // Mode 2: the networks of the plan's first kMlpLds MLP steps, copied to
    // LDS once per workgroup (a wave-wide inference then reads its weight
    // fragments from LDS, not from L2); later MLP steps run as CONIC.
    const MlpPacked* nets = nullptr;
    if constexpr (Mlp) {
        __shared__ MlpPacked lds_nets[kMlpLds];
        int slot = 0;
        for (int k = 0; k < plan.nstep && slot < kMlpLds; k++) {
            const bt_step* st = plan.steps + k;
            if (st->start != BT_START_MLP || !st->mlp) continue;
            const uint4* src = static_cast<const uint4*>(st->mlp);
            uint4* dst = reinterpret_cast<uint4*>(lds_nets + slot);
            for (int q = threadIdx.x; q < (int)(sizeof(MlpPacked) / 16); q += kBlock)
                dst[q] = src[q];
            slot++;
        }
        __syncthreads();
        nets = lds_nets;
    }
    const int64_t i0 = ray_index();
    if (!Mlp && i0 >= in.n) return;
    const bool live = i0 < in.n;
    const int64_t i = live ? i0 : in.n - 1;
    Ray r;
    r.fail = in.failed[i] != 0;
    r.vig = in.vignetted[i] != 0;
    r.x = in.x[i]; r.y = in.y[i]; r.z = in.z[i];
    r.vx = in.vx[i]; r.vy = in.vy[i]; r.vz = in.vz[i];
    r.t = in.t[i];
    // Refractive indices: a constant medium's is its record's (a scalar
    // load); the plan's first dispersive medium is evaluated once per ray and
    // kept; any other dispersive medium is evaluated at its refraction, with
    // the wavelength read again from memory (it is not held in registers).
    const int nmed = plan.nmedia;
    int cm = -1;                               // the cached medium (wave-uniform)
    for (int m = 0; m < nmed; m++)
        if (plan.media[m].kind != BT_MED_CONST) { cm = m; break; }
    double ncm = 0.0;
    if (need_w && cm >= 0 && !r.fail) ncm = medium_n(plan.media + cm, in.wavelength[i]);
    const int nstep = plan.nstep;
    int mslot = 0;                             // next LDS network (wave-uniform)
    for (int k = 0; k < nstep; k++) {
        if (!Mlp && r.fail && r.vig) break;    // nothing below can change this ray
        const bt_step* __restrict__ st = plan.steps + k;
        const bt_surface* surf = plan.surfaces + st->surface;
        const int action = st->action;
        double n2 = 0.0;
        if (action == BT_ACT_REFRACT && !r.fail) {
            const int mi = st->medium;
            const bt_medium* med = plan.media + mi;
            n2 = (med->kind == BT_MED_CONST) ? med->p[0]           // medium.cpp:30-32
                 : (mi == cm) ? ncm : medium_n(med, in.wavelength[i]);
        }
        double correction = 0.0;
        int start = Fast ? st->start : BT_START_REFERENCE;
        if (start == BT_START_MLP && (!Mlp || !st->mlp || mslot >= kMlpLds))
            start = BT_START_CONIC;
        if (Mlp && start == BT_START_MLP) {     // uniform: whole wave
            const MlpPacked* P = nets + mslot;
            mslot++;
            float f[kMlpIn];
            mlp_features<K>(st, surf, r, P, f);
            correction = (double)mlp_predict_wave(P, f);
        }
        apply_interaction<K>(action, st->dr, st->drot, surf, n2, r, niter, start, correction);
        const int olen = st->obsc_len;
        if (olen > 0 && !r.vig)
            r.vig = obsc_contains(plan.obscs + st->obsc, olen, r.x, r.y);
    }
    if (Mlp && !live) return;
    out.x[i] = r.x; out.y[i] = r.y; out.z[i] = r.z;
    out.vx[i] = r.vx; out.vy[i] = r.vy; out.vz[i] = r.vz;
    out.t[i] = r.t;
    out.vignetted[i] = r.vig;
    out.failed[i] = r.fail;
    if (out.wavelength != in.wavelength) out.wavelength[i] = in.wavelength[i];
    if (out.flux != in.flux) out.flux[i] = in.flux[i];
}

__global__ __launch_bounds__(kBlock) void k_sag(
    const bt_surface* __restrict__ s, const double* __restrict__ x,
    const double* __restrict__ y, double* __restrict__ out, int64_t n)
{
    const int64_t i = ray_index();
    if (i >= n) return;
    out[i] = surf_sag(s, x[i], y[i]);
}

__global__ __launch_bounds__(kBlock) void k_normal(
    const bt_surface* __restrict__ s, const double* __restrict__ x,
    const double* __restrict__ y, double* __restrict__ out, int64_t n)
{                                                    // pysrc/surface.cpp:12-21
    const int64_t i = ray_index();
    if (i >= n) return;
    double nx, ny, nz;
    surf_normal(s, x[i], y[i], nx, ny, nz);
    out[i] = nx; out[i+n] = ny; out[i+2*n] = nz;
}

__global__ __launch_bounds__(kBlock) void k_get_n(
    const bt_medium* __restrict__ m, const double* __restrict__ w,
    double* __restrict__ out, int64_t n)
{
    const int64_t i = ray_index();
    if (i >= n) return;
    out[i] = medium_n(m, w[i]);
}

__global__ __launch_bounds__(kBlock) void k_contains(
    const bt_obsc* __restrict__ o, int len, const double* __restrict__ x,
    const double* __restrict__ y, uint8_t* __restrict__ out, int64_t n)
{
    const int64_t i = ray_index();
    if (i >= n) return;
    out[i] = obsc_contains(o, len, x[i], y[i]) ? 1 : 0;
}

__global__ __launch_bounds__(kBlock) void k_check_sqrt(
    const double* __restrict__ x, double* __restrict__ out, int64_t n)
{
    const int64_t i = ray_index();
    if (i >= n) return;
    out[i] = cr_sqrt(x[i]);
    out[n + i] = cr_rsqrt(x[i]);
}

// ---------------------------------------------------------------------------
// Host helpers
// ---------------------------------------------------------------------------
bool valid_surface_kind(int k) { return k >= BT_PLANE && k <= BT_SUM; }

int step_launch(int action, const bt_surface* surface, const double* dr,
                const double* drot, const bt_medium* m2, bt_rays rays,
                int niter, const bt_coating* coating, const bt_obsc* obsc,
                int obsc_len, void* stream, const char* name, int live = 0,
                double min_flux = 0.0, int kinds = 0, int64_t* counts = nullptr)
{
    if (!surface || !dr || !drot) return fail(BT_E_NULL, std::string(name) + ": null argument");
    if (!rays_ok(rays)) return fail(BT_E_NULL, std::string(name) + ": incomplete ray columns");
    if (action == BT_ACT_REFRACT && !m2) return fail(BT_E_NULL, "bt_refract: null medium");
    if (obsc_len < 0 || (obsc_len > 0 && !obsc))
        return fail(BT_E_SIZE, std::string(name) + ": bad obscuration program");
    if (niter < 0) return fail(BT_E_SIZE, std::string(name) + ": negative niter");
    if (rays.n == 0) return 0;
    const double coat = !coating ? 1.0
                        : (action == BT_ACT_REFLECT ? coating->reflectivity
                                                    : coating->transmissivity);
    auto* c = reinterpret_cast<unsigned long long*>(counts);
    if (kinds & BT_PLAN_ANALYTIC)
        hipLaunchKernelGGL(k_step<kKindsAnalytic>, dim3(grid_for(rays.n)), dim3(kBlock), 0,
                           as_stream(stream), action, make_xform(dr, drot), surface, m2, obsc,
                           obsc_len, rays, niter, coating ? 1 : 0, coat, live, min_flux, c);
    else
        hipLaunchKernelGGL(k_step<kKindsAll>, dim3(grid_for(rays.n)), dim3(kBlock), 0,
                           as_stream(stream), action, make_xform(dr, drot), surface, m2, obsc,
                           obsc_len, rays, niter, coating ? 1 : 0, coat, live, min_flux, c);
    return check_launch(name);
}

}  // namespace bt

using namespace bt;

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

int bt_abi_version(void) { return BT_ABI_VERSION; }

const char* bt_last_error(void) { return g_last_error.c_str(); }

int bt_struct_sizes(int64_t* sizes, int n)
{
    const int64_t all[9] = {(int64_t)sizeof(bt_rays), (int64_t)sizeof(bt_surface),
                            (int64_t)sizeof(bt_medium), (int64_t)sizeof(bt_obsc),
                            (int64_t)sizeof(bt_step), (int64_t)sizeof(bt_plan),
                            (int64_t)sizeof(bt_psf_segment), (int64_t)sizeof(bt_raygen),
                            (int64_t)sizeof(bt_coating)};
    if (!sizes) return fail(BT_E_NULL, "bt_struct_sizes: null argument");
    for (int i = 0; i < n && i < 9; i++) sizes[i] = all[i];
    return 0;
}

namespace {
struct FieldOffset { const char* name; int64_t offset; };
#define BT_F(S, f) {#f, (int64_t)offsetof(S, f)}
const FieldOffset kRaysFields[] = {
    BT_F(bt_rays, x), BT_F(bt_rays, y), BT_F(bt_rays, z), BT_F(bt_rays, vx),
    BT_F(bt_rays, vy), BT_F(bt_rays, vz), BT_F(bt_rays, t), BT_F(bt_rays, wavelength),
    BT_F(bt_rays, flux), BT_F(bt_rays, vignetted), BT_F(bt_rays, failed), BT_F(bt_rays, n)};
const FieldOffset kSurfaceFields[] = {
    BT_F(bt_surface, kind), BT_F(bt_surface, nchild), BT_F(bt_surface, n1),
    BT_F(bt_surface, n2), BT_F(bt_surface, flags), BT_F(bt_surface, pad_),
    BT_F(bt_surface, p), BT_F(bt_surface, a0), BT_F(bt_surface, a1), BT_F(bt_surface, a2),
    BT_F(bt_surface, a3)};
const FieldOffset kMediumFields[] = {
    BT_F(bt_medium, kind), BT_F(bt_medium, n), BT_F(bt_medium, p), BT_F(bt_medium, a0),
    BT_F(bt_medium, a1)};
const FieldOffset kObscFields[] = {
    BT_F(bt_obsc, op), BT_F(bt_obsc, n), BT_F(bt_obsc, p), BT_F(bt_obsc, a0),
    BT_F(bt_obsc, a1)};
const FieldOffset kStepFields[] = {
    BT_F(bt_step, action), BT_F(bt_step, surface), BT_F(bt_step, medium),
    BT_F(bt_step, obsc), BT_F(bt_step, obsc_len), BT_F(bt_step, start), BT_F(bt_step, dr),
    BT_F(bt_step, drot), BT_F(bt_step, mlp)};
const FieldOffset kPlanFields[] = {
    BT_F(bt_plan, nstep), BT_F(bt_plan, flags), BT_F(bt_plan, nmedia), BT_F(bt_plan, pad_),
    BT_F(bt_plan, steps), BT_F(bt_plan, surfaces), BT_F(bt_plan, media),
    BT_F(bt_plan, obscs)};
const FieldOffset kSegmentFields[] = {
    BT_F(bt_psf_segment, begin), BT_F(bt_psf_segment, end), BT_F(bt_psf_segment, chief),
    BT_F(bt_psf_segment, time_ray), BT_F(bt_psf_segment, a1), BT_F(bt_psf_segment, a2)};
const FieldOffset kRaygenFields[] = {
    BT_F(bt_raygen, pattern), BT_F(bt_raygen, nring), BT_F(bt_raygen, nx),
    BT_F(bt_raygen, ny), BT_F(bt_raygen, fx), BT_F(bt_raygen, fy), BT_F(bt_raygen, lx),
    BT_F(bt_raygen, ly), BT_F(bt_raygen, lo), BT_F(bt_raygen, span),
    BT_F(bt_raygen, pcg_state), BT_F(bt_raygen, pcg_inc), BT_F(bt_raygen, ring_off),
    BT_F(bt_raygen, ring_rho), BT_F(bt_raygen, ring_dth), BT_F(bt_raygen, px),
    BT_F(bt_raygen, py), BT_F(bt_raygen, flux), BT_F(bt_raygen, stop),
    BT_F(bt_raygen, stop_dr), BT_F(bt_raygen, stop_drot), BT_F(bt_raygen, origin),
    BT_F(bt_raygen, rot), BT_F(bt_raygen, vv), BT_F(bt_raygen, wavelength),
    BT_F(bt_raygen, flux0)};
const FieldOffset kCoatingFields[] = {
    BT_F(bt_coating, reflectivity), BT_F(bt_coating, transmissivity)};
#undef BT_F
#define BT_TABLE(t) {t, (int)(sizeof(t) / sizeof(t[0]))}
const struct { const FieldOffset* fields; int count; } kStructFields[9] = {
    BT_TABLE(kRaysFields), BT_TABLE(kSurfaceFields), BT_TABLE(kMediumFields),
    BT_TABLE(kObscFields), BT_TABLE(kStepFields), BT_TABLE(kPlanFields),
    BT_TABLE(kSegmentFields), BT_TABLE(kRaygenFields), BT_TABLE(kCoatingFields)};
#undef BT_TABLE
}  // namespace

int bt_struct_field_count(int which)
{
    if (which < 0 || which >= 9) return fail(BT_E_SIZE, "bt_struct_field_count: no such struct");
    return kStructFields[which].count;
}

int64_t bt_struct_field_offset(int which, const char* field)
{
    if (which < 0 || which >= 9 || !field) return -1;
    for (int i = 0; i < kStructFields[which].count; i++)
        if (strcmp(kStructFields[which].fields[i].name, field) == 0)
            return kStructFields[which].fields[i].offset;
    return -1;
}

int bt_device_info(int* cu_count, int* arch_major, int* arch_minor)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail((int)e, "hipGetDevice failed");
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, dev);
    if (e != hipSuccess) return fail((int)e, "hipGetDeviceProperties failed");
    if (cu_count) *cu_count = prop.multiProcessorCount;
    if (arch_major) *arch_major = prop.major;
    if (arch_minor) *arch_minor = prop.minor;
    return 0;
}

int bt_apply_forward_transform(const double* dr, const double* drot,
                               bt_rays rays, void* stream)
{
    if (!dr || !drot || !rays_ok(rays)) return fail(BT_E_NULL, "bt_apply_forward_transform: null argument");
    if (rays.n == 0) return 0;
    hipLaunchKernelGGL(k_transform_rays<true>, dim3(grid_for(rays.n)), dim3(kBlock), 0,
                       as_stream(stream), make_xform(dr, drot), rays);
    return check_launch("bt_apply_forward_transform");
}

int bt_apply_reverse_transform(const double* dr, const double* drot,
                               bt_rays rays, void* stream)
{
    if (!dr || !drot || !rays_ok(rays)) return fail(BT_E_NULL, "bt_apply_reverse_transform: null argument");
    if (rays.n == 0) return 0;
    hipLaunchKernelGGL(k_transform_rays<false>, dim3(grid_for(rays.n)), dim3(kBlock), 0,
                       as_stream(stream), make_xform(dr, drot), rays);
    return check_launch("bt_apply_reverse_transform");
}

int bt_apply_forward_transform_arrays(const double* dr, const double* drot,
                                      double* x, double* y, double* z,
                                      int64_t n, void* stream)
{
    if (!dr || !drot || (n > 0 && (!x || !y || !z)))
        return fail(BT_E_NULL, "bt_apply_forward_transform_arrays: null argument");
    if (n < 0) return fail(BT_E_SIZE, "bt_apply_forward_transform_arrays: negative n");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_transform_arrays<true>, dim3(grid_for(n)), dim3(kBlock), 0,
                       as_stream(stream), make_xform(dr, drot), x, y, z, n);
    return check_launch("bt_apply_forward_transform_arrays");
}

int bt_apply_reverse_transform_arrays(const double* dr, const double* drot,
                                      double* x, double* y, double* z,
                                      int64_t n, void* stream)
{
    if (!dr || !drot || (n > 0 && (!x || !y || !z)))
        return fail(BT_E_NULL, "bt_apply_reverse_transform_arrays: null argument");
    if (n < 0) return fail(BT_E_SIZE, "bt_apply_reverse_transform_arrays: negative n");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_transform_arrays<false>, dim3(grid_for(n)), dim3(kBlock), 0,
                       as_stream(stream), make_xform(dr, drot), x, y, z, n);
    return check_launch("bt_apply_reverse_transform_arrays");
}

int bt_finish_parallel(const double* dr, const double* drot, const double* vv,
                       double* x, double* y, double* z, int64_t n, void* stream)
{
    if (!dr || !drot || !vv || (n > 0 && (!x || !y || !z)))
        return fail(BT_E_NULL, "bt_finish_parallel: null argument");
    if (n < 0) return fail(BT_E_SIZE, "bt_finish_parallel: negative n");
    if (n == 0) return 0;
    FinishArgs a;
    a.xf = make_xform(dr, drot);
    // batoid.cpp:44-46, evaluated once on the host as the reference does
    a.vl[0] = -vv[0]*drot[0] - vv[1]*drot[3] - vv[2]*drot[6];
    a.vl[1] = -vv[0]*drot[1] - vv[1]*drot[4] - vv[2]*drot[7];
    a.vl[2] = -vv[0]*drot[2] - vv[1]*drot[5] - vv[2]*drot[8];
    hipLaunchKernelGGL(k_finish_parallel, dim3(grid_for(n)), dim3(kBlock), 0,
                       as_stream(stream), a, x, y, z, n);
    return check_launch("bt_finish_parallel");
}

int bt_intersect(const bt_surface* surface, const double* dr,
                 const double* drot, bt_rays rays, int niter,
                 const bt_coating* coating, const bt_obsc* obsc, int obsc_len,
                 void* stream)
{
    return step_launch(BT_ACT_INTERSECT, surface, dr, drot, nullptr, rays, niter,
                       coating, obsc, obsc_len, stream, "bt_intersect");
}

int bt_reflect(const bt_surface* surface, const double* dr,
               const double* drot, bt_rays rays, int niter,
               const bt_coating* coating, const bt_obsc* obsc, int obsc_len,
               void* stream)
{
    return step_launch(BT_ACT_REFLECT, surface, dr, drot, nullptr, rays, niter,
                       coating, obsc, obsc_len, stream, "bt_reflect");
}

int bt_refract(const bt_surface* surface, const double* dr,
               const double* drot, const bt_medium* m2, bt_rays rays,
               int niter, const bt_coating* coating, const bt_obsc* obsc,
               int obsc_len, void* stream)
{
    return step_launch(BT_ACT_REFRACT, surface, dr, drot, m2, rays, niter,
                       coating, obsc, obsc_len, stream, "bt_refract");
}

int bt_interact_live(int action, const bt_surface* surface, const double* dr,
                     const double* drot, const bt_medium* m2, bt_rays rays, int niter,
                     const bt_coating* coating, const bt_obsc* obsc, int obsc_len,
                     double min_flux, int kinds, int64_t* counts, void* stream)
{
    if (action < BT_ACT_INTERSECT || action > BT_ACT_REFRACT)
        return fail(BT_E_SIZE, "bt_interact_live: unknown action");
    return step_launch(action, surface, dr, drot, m2, rays, niter, coating, obsc, obsc_len,
                       stream, "bt_interact_live", 1, min_flux, kinds, counts);
}

int bt_obscure(const bt_obsc* obsc, int obsc_len, bt_rays rays, void* stream)
{
    if (!obsc || obsc_len <= 0) return fail(BT_E_NULL, "bt_obscure: empty obscuration program");
    if (!rays_ok(rays)) return fail(BT_E_NULL, "bt_obscure: incomplete ray columns");
    if (rays.n == 0) return 0;
    hipLaunchKernelGGL(k_obscure, dim3(grid_for(rays.n)), dim3(kBlock), 0,
                       as_stream(stream), obsc, obsc_len, rays);
    return check_launch("bt_obscure");
}

int bt_trace_plan(bt_plan plan, bt_rays in, bt_rays out, int niter, void* stream)
{
    if (plan.nstep < 0) return fail(BT_E_PLAN, "bt_trace_plan: negative step count");
    if (plan.nstep > 0 && (!plan.steps || !plan.surfaces))
        return fail(BT_E_PLAN, "bt_trace_plan: plan tables missing");
    if (plan.nmedia < 0 || (plan.nmedia > 0 && !plan.media))
        return fail(BT_E_PLAN, "bt_trace_plan: bad media table");
    if (!rays_ok(in) || !rays_ok(out)) return fail(BT_E_NULL, "bt_trace_plan: incomplete ray columns");
    if (in.n != out.n) return fail(BT_E_SIZE, "bt_trace_plan: in/out sizes differ");
    if (niter < 0) return fail(BT_E_SIZE, "bt_trace_plan: negative niter");
    if (in.n == 0) return 0;
    const int need_w = (plan.flags & BT_PLAN_NO_REFRACT) ? 0 : 1;
    const int kinds = (plan.flags & BT_PLAN_ANALYTIC) ? kKindsAnalytic
                      : (plan.flags & BT_PLAN_SUM_APB) ? kKindsSumAPB
                      : (plan.flags & BT_PLAN_FLAT_SUMS) ? kKindsFlatSums : kKindsAll;
    const int mode = (plan.flags & BT_PLAN_MLP_START) ? 2 : (plan.flags & BT_PLAN_FAST_START) ? 1 : 0;
    const dim3 grid(grid_for(in.n)), block(kBlock);
    hipStream_t st = as_stream(stream);
#define BT_PLAN_LAUNCH(M, K) \
    hipLaunchKernelGGL((k_plan<M, K>), grid, block, 0, st, plan, in, out, niter, need_w)
    switch (mode * 4 + kinds) {
    case 0: BT_PLAN_LAUNCH(0, kKindsAll); break;
    case 1: BT_PLAN_LAUNCH(0, kKindsAnalytic); break;
    case 2: BT_PLAN_LAUNCH(0, kKindsSumAPB); break;
    case 3: BT_PLAN_LAUNCH(0, kKindsFlatSums); break;
    case 4: BT_PLAN_LAUNCH(1, kKindsAll); break;
    case 5: BT_PLAN_LAUNCH(1, kKindsAnalytic); break;
    case 6: BT_PLAN_LAUNCH(1, kKindsSumAPB); break;
    case 7: BT_PLAN_LAUNCH(1, kKindsFlatSums); break;
    case 8: BT_PLAN_LAUNCH(2, kKindsAll); break;
    case 9: BT_PLAN_LAUNCH(2, kKindsAnalytic); break;
    case 10: BT_PLAN_LAUNCH(2, kKindsSumAPB); break;
    default: BT_PLAN_LAUNCH(2, kKindsFlatSums); break;
    }
#undef BT_PLAN_LAUNCH
    return check_launch("bt_trace_plan");
}

int bt_surface_sag(const bt_surface* surface, const double* x, const double* y,
                   double* out, int64_t n, void* stream)
{
    if (!surface || (n > 0 && (!x || !y || !out))) return fail(BT_E_NULL, "bt_surface_sag: null argument");
    if (n < 0) return fail(BT_E_SIZE, "bt_surface_sag: negative n");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_sag, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream),
                       surface, x, y, out, n);
    return check_launch("bt_surface_sag");
}

int bt_surface_normal(const bt_surface* surface, const double* x,
                      const double* y, double* normal_out, int64_t n,
                      void* stream)
{
    if (!surface || (n > 0 && (!x || !y || !normal_out)))
        return fail(BT_E_NULL, "bt_surface_normal: null argument");
    if (n < 0) return fail(BT_E_SIZE, "bt_surface_normal: negative n");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_normal, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream),
                       surface, x, y, normal_out, n);
    return check_launch("bt_surface_normal");
}

int bt_medium_get_n(const bt_medium* medium, const double* wavelength,
                    double* out, int64_t n, void* stream)
{
    if (!medium || (n > 0 && (!wavelength || !out))) return fail(BT_E_NULL, "bt_medium_get_n: null argument");
    if (n < 0) return fail(BT_E_SIZE, "bt_medium_get_n: negative n");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_get_n, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream),
                       medium, wavelength, out, n);
    return check_launch("bt_medium_get_n");
}

int bt_obsc_contains(const bt_obsc* obsc, int obsc_len, const double* x,
                     const double* y, uint8_t* out, int64_t n, void* stream)
{
    if (!obsc || obsc_len <= 0 || (n > 0 && (!x || !y || !out)))
        return fail(BT_E_NULL, "bt_obsc_contains: null argument");
    if (n < 0) return fail(BT_E_SIZE, "bt_obsc_contains: negative n");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_contains, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream),
                       obsc, obsc_len, x, y, out, n);
    return check_launch("bt_obsc_contains");
}

int bt_check_sqrt(const double* x, double* out, int64_t n, void* stream)
{
    if (n > 0 && (!x || !out)) return fail(BT_E_NULL, "bt_check_sqrt: null argument");
    if (n < 0) return fail(BT_E_SIZE, "bt_check_sqrt: negative n");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_check_sqrt, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream),
                       x, out, n);
    return check_launch("bt_check_sqrt");
}

}  // extern "C"
