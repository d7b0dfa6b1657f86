// bt_split.hip -- ghost splitting: rSplit and the ray selection of traceSplit.
//
// bt_rsplit splits every ray that reaches a surface into its refracted and
// reflected parts in one pass (src/batoid.cpp:509-674): lane i intersects
// ray i with the reference's Newton iteration, then writes the reflected
// ray to the split buffer and the refracted ray back in place, with the
// coating's flux fractions.  The arithmetic is the reference's, operation by
// operation.
//
// traceSplit (batoid/optic.py:1244-1282) keeps, after every Interface, only
// the rays that are unvignetted and above a flux threshold.  Here that
// filter is an order-preserving stream compaction: per-block counts from
// wave ballots, one exclusive scan over the blocks, then a scatter that
// recomputes each ray's rank from the same ballots.
#include <hip/hip_runtime.h>
#include <string>

#include "bt_device.h"
#include "bt_common.h"

namespace bt {

// One ray of rSplit; returns false for a ray traceSplit's filter had
// already dropped (live mode), which is left untouched.  `lr` / `ls`: the
// refracted and split parts pass traceSplit's filter afterwards.
template <int K>
__device__ __forceinline__ void rsplit_ray(
    int64_t i, const Xform& xf, const bt_surface* __restrict__ surf,
    const bt_medium* __restrict__ m2, double reflect, double transmit,
    const bt_obsc* __restrict__ obsc, int olen, const bt_rays& rv, const bt_rays& sp,
    int niter, int live, double min_flux, bool& lr, bool& ls)
{                                                    // batoid.cpp:582-672
    lr = ls = false;
    // traceSplit's filter (optic.py:1262) has already dropped this ray: it
    // stays as it is, and its split part is only marked dropped too
    if (live && (rv.vignetted[i] != 0 || !(rv.flux[i] >= min_flux))) {
        if (sp.vignetted) sp.vignetted[i] = 1;
        return;
    }
    const double* dr = xf.dr;
    const double* drot = xf.drot;
    const double x0 = rv.x[i], y0 = rv.y[i], z0 = rv.z[i];
    const double vx0 = rv.vx[i], vy0 = rv.vy[i], vz0 = rv.vz[i];
    const double t0 = rv.t[i], w = rv.wavelength[i], f = rv.flux[i];
    const bool vig0 = rv.vignetted[i] != 0, fail0 = rv.failed[i] != 0;
    // the split ray starts as a copy of the input (trace.py:125)
    double sx = x0, sy = y0, sz = z0, svx = vx0, svy = vy0, svz = vz0, st = t0, sf = f;
    bool svig = vig0, sfail = fail0;
    // the refracted ray's final position, flag and flux, for the obscuration
    // and the filter
    double rx = x0, ry = y0, rf = f;
    bool rvig = vig0;
    if (!fail0) {
        double dx = x0-dr[0];
        double dy = y0-dr[1];
        double dz = z0-dr[2];
        double x = dx*drot[0] + dy*drot[3] + dz*drot[6];
        double y = dx*drot[1] + dy*drot[4] + dz*drot[7];
        double z = dx*drot[2] + dy*drot[5] + dz*drot[8];
        double vx = vx0*drot[0] + vy0*drot[3] + vz0*drot[6];
        double vy = vx0*drot[1] + vy0*drot[4] + vz0*drot[7];
        double vz = vx0*drot[2] + vy0*drot[5] + vz0*drot[8];
        double t = t0;
        double dt = 0.0;
        NormalMemo memo;
        memo.ok = false;
        const bool hit = surf_tti<K>(surf, x, y, z, vx, vy, vz, dt, niter, &memo);
        if (hit) {
            x += vx * dt;
            y += vy * dt;
            z += vz * dt;
            t += dt;
            double n1 = vx*vx;
            n1 += vy*vy;
            n1 += vz*vz;
            n1 = cr_rsqrt(n1);
            double nvx = vx*n1;
            double nvy = vy*n1;
            double nvz = vz*n1;
            double nx, ny, nz;
            if (memo.ok && same_bits(memo.dt, dt)) {  // Newton's last normal, same point
                nx = memo.nx; ny = memo.ny; nz = memo.nz;
            } else {
                surf_normal<K>(surf, x, y, nx, ny, nz);
            }
            double alpha = nvx*nx;
            alpha += nvy*ny;
            alpha += nvz*nz;
            if (alpha > 0) {
                nx *= -1;
                ny *= -1;
                nz *= -1;
                alpha *= -1;
            }
            // reflection
            sx = x; sy = y; sz = z;
            svx = vx - 2*alpha*nx/n1;
            svy = vy - 2*alpha*ny/n1;
            svz = vz - 2*alpha*nz/n1;
            st = t;
            sf = f*reflect;
            // refraction
            const double n2 = medium_n(m2, w);
            const double eta = n1/n2;
            const double sinsqr = eta*eta*(1-alpha*alpha);
            rv.x[i] = x;
            rv.y[i] = y;
            rv.z[i] = z;
            rv.t[i] = t;
            rx = x;
            ry = y;
            if (sinsqr > 1.0) {                      // total internal reflection
                const double nan = __builtin_nan("");
                rv.vx[i] = nan; rv.vy[i] = nan; rv.vz[i] = nan;
                rv.failed[i] = 1;
                rv.vignetted[i] = 1;
                rvig = true;
            } else {
                const double nfactor = eta*alpha + cr_sqrt(1-sinsqr);
                double ux = eta*nvx - nfactor*nx;
                double uy = eta*nvy - nfactor*ny;
                double uz = eta*nvz - nfactor*nz;
                rv.vx[i] = ux / n2;
                rv.vy[i] = uy / n2;
                rv.vz[i] = uz / n2;
            }
            rf = f*transmit;
            rv.flux[i] = rf;
        } else {
            rv.failed[i] = 1;
            rv.vignetted[i] = 1;
            rvig = true;
            svig = true;
            sfail = true;
        }
    }
    // the Interface's obscuration on both parts (optic.py:491-494), every
    // ray included, as batoid.cpp:151-172
    if (olen > 0) {
        if (!rvig && obsc_contains(obsc, olen, rx, ry)) {
            rv.vignetted[i] = 1;
            rvig = true;
        }
        if (!svig) svig = obsc_contains(obsc, olen, sx, sy);
    }
    lr = !rvig && rf >= min_flux;
    ls = !svig && sf >= min_flux;
    if (!sp.flux) return;                           // no split part asked for
    sp.flux[i] = sf;
    sp.vignetted[i] = svig;
    sp.failed[i] = sfail;
    // live mode: a split ray the filter drops (vignetted, or its flux below
    // min_flux -- every one once the ray's flux times the reflectivity is)
    // is never copied out or traced further, and every later pass reads
    // only its flux and flags
    if (live && !ls) return;
    sp.x[i] = sx; sp.y[i] = sy; sp.z[i] = sz;
    sp.vx[i] = svx; sp.vy[i] = svy; sp.vz[i] = svz;
    sp.t[i] = st;
    sp.wavelength[i] = w;
}

// `counts` (live mode, may be null): the numbers of refracted and split
// rays that pass traceSplit's filter, added up by wave ballots into
// kCountSlots partial sums each (the wave's block picks the slot), so that
// no one address takes every wave's atomic.
constexpr int kCountSlots = 64;

template <int K>
__global__ __launch_bounds__(kBlock) void k_rsplit(
    Xform xf, const bt_surface* __restrict__ surf, const bt_medium* __restrict__ m2,
    double reflect, double transmit, const bt_obsc* __restrict__ obsc, int olen, bt_rays rv,
    bt_rays sp, int niter, int live, double min_flux, unsigned long long* __restrict__ counts)
{
    const int64_t i = (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    bool lr = false, ls = false;
    if (i < rv.n)
        rsplit_ray<K>(i, xf, surf, m2, reflect, transmit, obsc, olen, rv, sp, niter, live,
                      min_flux, lr, ls);
    if (counts) {
        const uint64_t mr = __ballot(lr), ms = __ballot(ls);
        if ((threadIdx.x & 63) == 0) {
            const int slot = (int)(blockIdx.x % kCountSlots);
            if (mr) atomicAdd(counts + slot, (unsigned long long)__popcll(mr));
            if (ms) atomicAdd(counts + kCountSlots + slot, (unsigned long long)__popcll(ms));
        }
    }
}

// ---------------------------------------------------------------------------
// Order-preserving selection: block b covers rays [256 b, 256 b + 256).
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool selected(const bt_rays& in, int64_t i, double min_flux)
{                                                    // optic.py:1262
    return i < in.n && in.vignetted[i] == 0 && in.flux[i] >= min_flux;
}

// Rank of this lane among the block's selected rays, and the block total.
__device__ __forceinline__ int block_rank(bool keep, int& total)
{
    __shared__ int wave_count[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint64_t m = __ballot(keep);
    if (lane == 0) wave_count[wave] = __popcll(m);
    __syncthreads();
    int base = 0;
    total = 0;
    for (int k = 0; k < kBlock / 64; k++) {
        if (k < wave) base += wave_count[k];
        total += wave_count[k];
    }
    return base + __popcll(m & ((1ull << lane) - 1ull));
}

__global__ __launch_bounds__(kBlock) void k_select_count(bt_rays in, double min_flux,
                                                         int64_t* __restrict__ counts)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    int total;
    block_rank(selected(in, i, min_flux), total);
    if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

// Exclusive scan of the block counts by one workgroup, in chunks of
// kBlock * kScanPer counts taken in order: each thread sums kScanPer
// consecutive counts, the thread sums are scanned within each wave by
// shuffles and across the four waves through LDS (two barriers per chunk).
constexpr int kScanPer = 4;

__global__ __launch_bounds__(kBlock) void k_select_scan(const int64_t* __restrict__ counts,
                                                        int64_t nblocks,
                                                        int64_t* __restrict__ offsets,
                                                        int64_t* __restrict__ total)
{
    __shared__ int64_t wave_sum[kBlock / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int64_t carry = 0;
    for (int64_t base = 0; base < nblocks; base += (int64_t)kBlock * kScanPer) {
        const int64_t b0 = base + (int64_t)threadIdx.x * kScanPer;
        int64_t c[kScanPer];
        int64_t sum = 0;
        for (int k = 0; k < kScanPer; k++) {
            c[k] = b0 + k < nblocks ? counts[b0 + k] : 0;
            sum += c[k];
        }
        int64_t incl = sum;                          // inclusive scan within the wave
        for (int d = 1; d < 64; d <<= 1) {
            const int64_t v = __shfl_up(incl, d, 64);
            if (lane >= d) incl += v;
        }
        if (lane == 63) wave_sum[wave] = incl;
        __syncthreads();
        int64_t before = carry, chunk = 0;
        for (int w = 0; w < kBlock / 64; w++) {
            if (w < wave) before += wave_sum[w];
            chunk += wave_sum[w];
        }
        __syncthreads();                             // wave_sum is rewritten next chunk
        int64_t run = before + incl - sum;
        for (int k = 0; k < kScanPer; k++) {
            if (b0 + k < nblocks) offsets[b0 + k] = run;
            run += c[k];
        }
        carry += chunk;
    }
    if (threadIdx.x == 0) *total = carry;
}

__global__ __launch_bounds__(kBlock) void k_select_scatter(bt_rays in, double min_flux,
                                                           const int64_t* __restrict__ offsets,
                                                           bt_rays out)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool keep = selected(in, i, min_flux);
    int total;
    const int rank = block_rank(keep, total);
    if (!keep) return;
    const int64_t o = offsets[blockIdx.x] + rank;
    if (o >= out.n) return;                          // out smaller than the count
    out.x[o] = in.x[i]; out.y[o] = in.y[i]; out.z[o] = in.z[i];
    out.vx[o] = in.vx[i]; out.vy[o] = in.vy[i]; out.vz[o] = in.vz[i];
    out.t[o] = in.t[i];
    out.wavelength[o] = in.wavelength[i];
    out.flux[o] = in.flux[i];
    out.vignetted[o] = in.vignetted[i];
    out.failed[o] = in.failed[i];
}

int64_t select_blocks(int64_t n) { return (n + kBlock - 1) / kBlock; }

}  // namespace bt

using namespace bt;

extern "C" {

namespace {
int rsplit_launch(const bt_surface* surface, const double* dr, const double* drot,
                  const bt_medium* m2, const bt_coating* coating, const bt_obsc* obsc,
                  int obsc_len, bt_rays rays, bt_rays split, int niter, int live,
                  double min_flux, int kinds, int64_t* counts, void* stream, const char* name)
{
    const std::string n(name);
    if (!surface || !dr || !drot || !m2 || !coating)
        return fail(BT_E_NULL, n + ": null argument");
    if (obsc_len < 0 || (obsc_len > 0 && !obsc))
        return fail(BT_E_SIZE, n + ": bad obscuration program");
    if (niter < 0) return fail(BT_E_SIZE, n + ": negative niter");
    // live mode may ask for no split part: every split column NULL, n = 0
    const bool no_split = live && split.n == 0 && !split.x && !split.y && !split.z &&
                          !split.vx && !split.vy && !split.vz && !split.t &&
                          !split.wavelength && !split.flux && !split.vignetted && !split.failed;
    if (rays.n < 0 || (!no_split && split.n < rays.n))
        return fail(BT_E_SIZE, n + ": split holds fewer rays");
    if (rays.n == 0) return 0;
    if (!rays_ok(rays) || (!no_split && !rays_ok(split)))
        return fail(BT_E_NULL, n + ": incomplete ray columns");
    if (rays.n > (int64_t)0x7fffffff * kBlock)
        return fail(BT_E_SIZE, n + ": too many rays");
    auto* c = reinterpret_cast<unsigned long long*>(counts);
    if (kinds & BT_PLAN_ANALYTIC)
        hipLaunchKernelGGL(k_rsplit<kKindsAnalytic>, dim3(grid_for(rays.n)), dim3(kBlock), 0,
                           as_stream(stream), make_xform(dr, drot), surface, m2,
                           coating->reflectivity, coating->transmissivity, obsc, obsc_len, rays,
                           split, niter, live, min_flux, c);
    else
        hipLaunchKernelGGL(k_rsplit<kKindsAll>, dim3(grid_for(rays.n)), dim3(kBlock), 0,
                           as_stream(stream), make_xform(dr, drot), surface, m2,
                           coating->reflectivity, coating->transmissivity, obsc, obsc_len, rays,
                           split, niter, live, min_flux, c);
    return check_launch(name);
}
}  // namespace

int bt_rsplit(const bt_surface* surface, const double* dr, const double* drot,
              const bt_medium* m2, const bt_coating* coating, const bt_obsc* obsc,
              int obsc_len, bt_rays rays, bt_rays split, int niter, void* stream)
{
    return rsplit_launch(surface, dr, drot, m2, coating, obsc, obsc_len, rays, split, niter, 0,
                         0.0, 0, nullptr, stream, "bt_rsplit");
}

int bt_rsplit_live(const bt_surface* surface, const double* dr, const double* drot,
                   const bt_medium* m2, const bt_coating* coating, const bt_obsc* obsc,
                   int obsc_len, bt_rays rays, bt_rays split, double min_flux, int niter,
                   int kinds, int64_t* counts, void* stream)
{
    return rsplit_launch(surface, dr, drot, m2, coating, obsc, obsc_len, rays, split, niter, 1,
                         min_flux, kinds, counts, stream, "bt_rsplit_live");
}

int64_t bt_select_workspace_bytes(int64_t n)
{
    if (n < 0) return -1;
    return 2 * (select_blocks(n) + 1) * (int64_t)sizeof(int64_t);
}

int bt_select_count(bt_rays in, double min_flux, int64_t* count, void* workspace,
                    int64_t workspace_bytes, void* stream)
{
    if ((in.n > 0 && !rays_ok(in)) || in.n < 0 || !count || !workspace)
        return fail(BT_E_NULL, "bt_select_count: null argument");
    if (workspace_bytes < bt_select_workspace_bytes(in.n))
        return fail(BT_E_SIZE, "bt_select_count: workspace too small");
    const int64_t nb = select_blocks(in.n);
    int64_t* counts = static_cast<int64_t*>(workspace);
    int64_t* offsets = counts + nb + 1;
    hipStream_t s = as_stream(stream);
    if (nb > 0) {
        if (nb > 0x7fffffff) return fail(BT_E_SIZE, "bt_select_count: too many rays");
        hipLaunchKernelGGL(k_select_count, dim3((unsigned)nb), dim3(kBlock), 0, s, in,
                           min_flux, counts);
        const int rc = check_launch("bt_select_count");
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_select_scan, dim3(1), dim3(kBlock), 0, s, (const int64_t*)counts, nb,
                       offsets, count);
    return check_launch("bt_select_count");
}

int bt_select_scatter(bt_rays in, double min_flux, bt_rays out, const void* workspace,
                      int64_t workspace_bytes, void* stream)
{
    if ((in.n > 0 && !rays_ok(in)) || in.n < 0 || !workspace ||
        (out.n > 0 && !rays_ok(out)))
        return fail(BT_E_NULL, "bt_select_scatter: null argument");
    if (workspace_bytes < bt_select_workspace_bytes(in.n))
        return fail(BT_E_SIZE, "bt_select_scatter: workspace too small");
    const int64_t nb = select_blocks(in.n);
    if (nb == 0 || out.n == 0) return 0;
    const int64_t* offsets = static_cast<const int64_t*>(workspace) + nb + 1;
    hipLaunchKernelGGL(k_select_scatter, dim3((unsigned)nb), dim3(kBlock), 0,
                       as_stream(stream), in, min_flux, offsets, out);
    return check_launch("bt_select_scatter");
}

}  // extern "C"
