// bt_field.hip -- RayVector field operations and the FP64 direct Huygens sum.
//
// positionAtTime / propagateInPlace / phase / amplitude / sumAmplitude of
// src/rayVector.cpp:26-222, and the huygensPSF pixel loop
// (batoid/analysis.py:296-303) evaluated on the GPU in FP64 with the
// reference's per-(pixel, ray) phase arithmetic.  Reductions are two-pass
// with a fixed combination order, so results are run-to-run deterministic.
#include <hip/hip_runtime.h>
#include <string>

#include "bt_device.h"
#include "bt_common.h"

namespace bt {

constexpr double kPi = 3.14159265358979323846;   // rayVector.cpp:88
constexpr int kSumBlocks = 1024;

__device__ __forceinline__ int64_t lin_index()
{
    return (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
}

// phi = 2 pi/lambda * (((p - r).v)/|v|^2 - (t_p - t))   rayVector.cpp:113-120
__device__ __forceinline__ double ray_phase(double px, double py, double pz,
                                            double pt, double x, double y,
                                            double z, double vx, double vy,
                                            double vz, double t, double w)
{
    double v2 = vx*vx + vy*vy + vz*vz;
    double out = (px-x)*vx;
    out += (py-y)*vy;
    out += (pz-z)*vz;
    out /= v2;
    out -= pt-t;
    out *= 2 * kPi / w;
    return out;
}

__global__ __launch_bounds__(kBlock) void k_position_at_time(
    bt_rays rv, double t, double* __restrict__ xo, double* __restrict__ yo,
    double* __restrict__ zo)
{                                                    // rayVector.cpp:49-53
    const int64_t i = lin_index();
    if (i >= rv.n) return;
    const double dt = t-rv.t[i];
    xo[i] = rv.x[i] + rv.vx[i] * dt;
    yo[i] = rv.y[i] + rv.vy[i] * dt;
    zo[i] = rv.z[i] + rv.vz[i] * dt;
}

__global__ __launch_bounds__(kBlock) void k_propagate(bt_rays rv, double t)
{                                                    // rayVector.cpp:76-81
    const int64_t i = lin_index();
    if (i >= rv.n) return;
    const double dt = t - rv.t[i];
    rv.x[i] += rv.vx[i] * dt;
    rv.y[i] += rv.vy[i] * dt;
    rv.z[i] += rv.vz[i] * dt;
    rv.t[i] = t;
}

__global__ __launch_bounds__(kBlock) void k_phase(
    bt_rays rv, double px, double py, double pz, double pt,
    double* __restrict__ out, int amplitude)
{                                                    // rayVector.cpp:84-166
    const int64_t i = lin_index();
    if (i >= rv.n) return;
    double ph = ray_phase(px, py, pz, pt, rv.x[i], rv.y[i], rv.z[i], rv.vx[i],
                          rv.vy[i], rv.vz[i], rv.t[i], rv.wavelength[i]);
    if (!amplitude) {
        out[i] = ph;
    } else {
        double s, c;
        sincos(ph, &s, &c);
        out[2*i] = c;
        out[2*i+1] = s;
    }
}

// Block-level FP64 pair reduction in a fixed tree order.
__device__ __forceinline__ void block_sum2(double& re, double& im, double* sre,
                                           double* sim)
{
    const int tid = threadIdx.x;
    sre[tid] = re; sim[tid] = im;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if (tid < s) { sre[tid] += sre[tid+s]; sim[tid] += sim[tid+s]; }
        __syncthreads();
    }
    re = sre[0]; im = sim[0];
}

__global__ __launch_bounds__(kBlock) void k_sum_amplitude_partial(
    bt_rays rv, double px, double py, double pz, double pt, int ignore_vig,
    double* __restrict__ partial)
{                                                    // rayVector.cpp:202-220
    __shared__ double sre[kBlock], sim[kBlock];
    double re = 0.0, im = 0.0;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < rv.n; i += stride) {
        if (!rv.failed[i] && !(ignore_vig && rv.vignetted[i])) {
            double ph = ray_phase(px, py, pz, pt, rv.x[i], rv.y[i], rv.z[i], rv.vx[i],
                                  rv.vy[i], rv.vz[i], rv.t[i], rv.wavelength[i]);
            double s, c;
            sincos(ph, &s, &c);
            re += c*rv.flux[i];
            im += s*rv.flux[i];
        }
    }
    block_sum2(re, im, sre, sim);
    if (threadIdx.x == 0) { partial[2*blockIdx.x] = re; partial[2*blockIdx.x+1] = im; }
}

__global__ __launch_bounds__(kBlock) void k_sum_partials(
    const double* __restrict__ partial, int nparts, int stride_pairs,
    int nout, double* __restrict__ out)
{
    // out[o] = sum_k partial[k*stride_pairs + o], k < nparts, fixed order
    __shared__ double sre[kBlock], sim[kBlock];
    const int o = blockIdx.x;
    if (o >= nout) return;
    double re = 0.0, im = 0.0;
    for (int k = threadIdx.x; k < nparts; k += blockDim.x) {
        re += partial[2*((int64_t)k*stride_pairs + o)];
        im += partial[2*((int64_t)k*stride_pairs + o)+1];
    }
    block_sum2(re, im, sre, sim);
    if (threadIdx.x == 0) { out[2*o] = re; out[2*o+1] = im; }
}

// Direct Huygens sum: one thread per pixel, rays staged through LDS in
// chunks; grid.y splits the rays, partial images reduced by k_sum_partials.
constexpr int kDirectChunk = 256;

__global__ __launch_bounds__(kBlock) void k_huygens_direct(
    bt_rays rv, double p0x, double p0y, double a1x, double a1y, double a2x,
    double a2y, int nx, int ny, double pt, int ignore_vig, int64_t rays_per_split,
    double* __restrict__ partial)
{
    __shared__ double s_x[kDirectChunk], s_y[kDirectChunk], s_z[kDirectChunk];
    __shared__ double s_vx[kDirectChunk], s_vy[kDirectChunk], s_vz[kDirectChunk];
    __shared__ double s_t[kDirectChunk], s_w[kDirectChunk], s_f[kDirectChunk];

    const int npix = nx * ny;
    const int pix = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = pix / nx, i = pix - (pix / nx) * nx;
    // lattice.py:6-14 coordinates, transposed as at analysis.py:296-297
    const double ci = (double)(i - nx/2), cj = (double)(j - ny/2);
    const double px = ci*a1x + cj*a2x + p0x;
    const double py = ci*a1y + cj*a2y + p0y;

    const int64_t r0 = (int64_t)blockIdx.y * rays_per_split;
    int64_t r1 = r0 + rays_per_split;
    if (r1 > rv.n) r1 = rv.n;
    double re = 0.0, im = 0.0;
    for (int64_t c0 = r0; c0 < r1; c0 += kDirectChunk) {
        __syncthreads();
        const int64_t r = c0 + threadIdx.x;
        if (threadIdx.x < kDirectChunk) {
            bool use = r < r1 && !rv.failed[r] && !(ignore_vig && rv.vignetted[r]);
            s_x[threadIdx.x] = use ? rv.x[r] : 0.0;
            s_y[threadIdx.x] = use ? rv.y[r] : 0.0;
            s_z[threadIdx.x] = use ? rv.z[r] : 0.0;
            s_vx[threadIdx.x] = use ? rv.vx[r] : 0.0;
            s_vy[threadIdx.x] = use ? rv.vy[r] : 0.0;
            s_vz[threadIdx.x] = use ? rv.vz[r] : 1.0;
            s_t[threadIdx.x] = use ? rv.t[r] : 0.0;
            s_w[threadIdx.x] = use ? rv.wavelength[r] : 1.0;
            s_f[threadIdx.x] = use ? rv.flux[r] : 0.0;
        }
        __syncthreads();
        const int cn = (int)((r1 - c0) < kDirectChunk ? (r1 - c0) : kDirectChunk);
        if (pix < npix) {
            for (int k = 0; k < cn; k++) {
                const double f = s_f[k];
                if (f == 0.0) continue;
                double ph = ray_phase(px, py, 0.0, pt, s_x[k], s_y[k], s_z[k],
                                      s_vx[k], s_vy[k], s_vz[k], s_t[k], s_w[k]);
                double s, c;
                sincos(ph, &s, &c);
                re += c*f;
                im += s*f;
            }
        }
    }
    if (pix < npix) {
        partial[2*((int64_t)blockIdx.y*npix + pix)] = re;
        partial[2*((int64_t)blockIdx.y*npix + pix)+1] = im;
    }
}

}  // namespace bt

using namespace bt;

extern "C" {

int bt_position_at_time(bt_rays rays, double t, double* xout, double* yout,
                        double* zout, void* stream)
{
    if (!rays_ok(rays) || (rays.n > 0 && (!xout || !yout || !zout)))
        return fail(BT_E_NULL, "bt_position_at_time: null argument");
    if (rays.n == 0) return 0;
    hipLaunchKernelGGL(k_position_at_time, dim3(grid_for(rays.n)), dim3(kBlock), 0,
                       as_stream(stream), rays, t, xout, yout, zout);
    return check_launch("bt_position_at_time");
}

int bt_propagate_in_place(bt_rays rays, double t, void* stream)
{
    if (!rays_ok(rays)) return fail(BT_E_NULL, "bt_propagate_in_place: null argument");
    if (rays.n == 0) return 0;
    hipLaunchKernelGGL(k_propagate, dim3(grid_for(rays.n)), dim3(kBlock), 0,
                       as_stream(stream), rays, t);
    return check_launch("bt_propagate_in_place");
}

int bt_phase(bt_rays rays, double x, double y, double z, double t,
             double* out, void* stream)
{
    if (!rays_ok(rays) || (rays.n > 0 && !out)) return fail(BT_E_NULL, "bt_phase: null argument");
    if (rays.n == 0) return 0;
    hipLaunchKernelGGL(k_phase, dim3(grid_for(rays.n)), dim3(kBlock), 0,
                       as_stream(stream), rays, x, y, z, t, out, 0);
    return check_launch("bt_phase");
}

int bt_amplitude(bt_rays rays, double x, double y, double z, double t,
                 double* amp_out, void* stream)
{
    if (!rays_ok(rays) || (rays.n > 0 && !amp_out)) return fail(BT_E_NULL, "bt_amplitude: null argument");
    if (rays.n == 0) return 0;
    hipLaunchKernelGGL(k_phase, dim3(grid_for(rays.n)), dim3(kBlock), 0,
                       as_stream(stream), rays, x, y, z, t, amp_out, 1);
    return check_launch("bt_amplitude");
}

int64_t bt_sum_amplitude_workspace_bytes(int64_t n)
{
    (void)n;
    return (int64_t)kSumBlocks * 2 * sizeof(double);
}

int bt_sum_amplitude(bt_rays rays, double x, double y, double z, double t,
                     int ignore_vignetted, double* out, void* workspace,
                     void* stream)
{
    if (!rays_ok(rays) || !out || !workspace) return fail(BT_E_NULL, "bt_sum_amplitude: null argument");
    int64_t g64 = (rays.n + kBlock - 1) / kBlock;
    const int g = (int)(g64 < 1 ? 1 : (g64 > kSumBlocks ? kSumBlocks : g64));
    double* partial = static_cast<double*>(workspace);
    hipLaunchKernelGGL(k_sum_amplitude_partial, dim3(g), dim3(kBlock), 0, as_stream(stream),
                       rays, x, y, z, t, ignore_vignetted, partial);
    int rc = check_launch("bt_sum_amplitude");
    if (rc) return rc;
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(kBlock), 0, as_stream(stream),
                       partial, g, 1, 1, out);
    return check_launch("bt_sum_amplitude(final)");
}

}  // extern "C"

namespace bt {

// Number of ray splits for the direct Huygens sum: enough blocks to cover
// the 256 CUs several times over.
int direct_splits(int64_t n, int npix)
{
    const int pix_blocks = (npix + kBlock - 1) / kBlock;
    int splits = (2048 + pix_blocks - 1) / pix_blocks;
    int64_t max_by_rays = (n + kDirectChunk - 1) / kDirectChunk;
    if (max_by_rays < 1) max_by_rays = 1;
    if (splits > max_by_rays) splits = (int)max_by_rays;
    if (splits > 256) splits = 256;
    if (splits < 1) splits = 1;
    return splits;
}

}  // namespace bt

extern "C" int64_t bt_huygens_direct_workspace_bytes(int64_t n, int nx, int ny)
{
    const int npix = nx * ny;
    return (int64_t)direct_splits(n, npix) * npix * 2 * (int64_t)sizeof(double);
}

extern "C" int bt_huygens_direct(bt_rays rays, const double* p0, const double* a1,
                                 const double* a2, int nx, int ny, double t,
                                 int ignore_vignetted, double* amp_out,
                                 void* workspace, int64_t workspace_bytes,
                                 void* stream)
{
    if (!rays_ok(rays) || !p0 || !a1 || !a2 || !amp_out || !workspace)
        return fail(BT_E_NULL, "bt_huygens_direct: null argument");
    if (nx <= 0 || ny <= 0 || (int64_t)nx * ny > (1 << 26))
        return fail(BT_E_SIZE, "bt_huygens_direct: bad image size");
    const int npix = nx * ny;
    const int splits = direct_splits(rays.n, npix);
    if (workspace_bytes < (int64_t)splits * npix * 2 * (int64_t)sizeof(double))
        return fail(BT_E_SIZE, "bt_huygens_direct: workspace too small");
    int64_t per = (rays.n + splits - 1) / splits;
    per = ((per + kDirectChunk - 1) / kDirectChunk) * kDirectChunk;
    if (per < kDirectChunk) per = kDirectChunk;
    double* partial = static_cast<double*>(workspace);
    dim3 grid((npix + kBlock - 1) / kBlock, splits);
    hipLaunchKernelGGL(k_huygens_direct, grid, dim3(kBlock), 0, as_stream(stream), rays,
                       p0[0], p0[1], a1[0], a1[1], a2[0], a2[1], nx, ny, t,
                       ignore_vignetted, per, partial);
    int rc = check_launch("bt_huygens_direct");
    if (rc) return rc;
    hipLaunchKernelGGL(k_sum_partials, dim3(npix), dim3(kBlock), 0, as_stream(stream),
                       partial, splits, npix, npix, amp_out);
    return check_launch("bt_huygens_direct(final)");
}
