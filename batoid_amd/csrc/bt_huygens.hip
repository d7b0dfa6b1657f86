// bt_huygens.hip -- the Huygens PSF sum as a complex GEMM on MFMA.
//
// The reference evaluates, for every pixel p of the image lattice,
//   A(p) = sum_r flux_r exp(i phi_r(p)),
//   phi_r(p) = 2pi/lambda_r (((p - x_r).v_r)/|v_r|^2 - (t - t_r))
// (rayVector.cpp:169-222, one sumAmplitude per pixel at analysis.py:302-303).
// The phase is affine in p, and the pixels form a lattice
//   p(i, j) = p0 + ci a1 + cj a2,  ci = i - nx/2, cj = j - ny/2,
// so  phi_r(p) = phi_r(p0) + ci alpha_r + cj beta_r  and
//   A[j][i] = sum_r  c_r e^{i ci alpha_r} e^{i cj beta_r},  c_r = flux e^{i phi_r(p0)}
// is a complex GEMM  U[i][r] V[r][j]  with M = nx, N = ny, K = rays.
//
// Precision plan:
//  * per ray, in FP64: phi_r(p0) with the reference's formula, reduced to
//    revolutions mod 1; alpha, beta in revolutions per pixel step, each split
//    into hi (few enough significant bits that ci * hi is exact in FP32) and
//    lo parts;
//  * per (pixel, ray), in FP32: theta = fract(ci*hi) + ci*lo + phi0, then
//    sin(2 pi theta) on the transcendental unit -- no FP32 rounding of a large
//    phase ever happens;
//  * the GEMM runs on MFMA: v_mfma_f32_32x32x2_f32 (FP32 operands), or the
//    split-BF16 form x = hi + lo with three v_mfma_f32_32x32x16_bf16
//    products (hi*hi + hi*lo + lo*hi), FP32 accumulation either way;
//  * partial images of the K splits are summed in FP64 in a fixed order.
//
// Complex -> real: the output rows are (i, re|im) pairs and the K index
// (ray, re|im), with A'[2i][2r] = Re U, A'[2i][2r+1] = -Im U,
// A'[2i+1][2r] = Im U, A'[2i+1][2r+1] = Re U, B'[2r][j] = Re V,
// B'[2r+1][j] = Im V.  Operands are generated directly in the MFMA lane
// layout; each wave owns a 64 (i) x 64 (j) complex output tile (eight
// 32 x 32 accumulator blocks) and a contiguous range of rays.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "bt_device.h"
#include "bt_common.h"

extern "C" int64_t bt_huygens_direct_workspace_bytes(int64_t n, int nx, int ny);

namespace bt {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr double kTwoPi = 2 * 3.14159265358979323846;
constexpr int kTile = 128;          // workgroup tile (complex) in i and j
constexpr int kRayStride = 8;       // floats per prepared ray
constexpr int kRotStride = 8;       // floats per ray of row-block rotations
constexpr int kRayGroup = 8;        // rays per bf16 MFMA step; splits align to it
constexpr int kScanThreads = 1024;

struct PsfGeom {
    double p0x, p0y, a1x, a1y, a2x, a2y, t;
    int nx, ny;
    int keep_a, keep_b;             // significant bits kept in the hi parts
};

// Keep the leading `keep` significant bits of x (1 <= keep <= 24).
__device__ __forceinline__ float hi_bits(float x, int keep)
{
    const uint32_t mask = ~((1u << (24 - keep)) - 1u);
    return __uint_as_float(__float_as_uint(x) & mask);
}

__device__ __forceinline__ bool psf_valid(const bt_rays& rv, int64_t i, int ignore_vig)
{
    return !rv.failed[i] && !(ignore_vig && rv.vignetted[i]);
}

__global__ __launch_bounds__(kBlock) void k_psf_count(bt_rays rv, int ignore_vig,
                                                       int* __restrict__ counts)
{
    __shared__ int s[kBlock / 64];
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool v = i < rv.n && psf_valid(rv, i, ignore_vig);
    const uint64_t m = __ballot(v);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = __popcll(m);
    __syncthreads();
    if (threadIdx.x == 0) {
        int c = 0;
        for (int w = 0; w < kBlock / 64; w++) c += s[w];
        counts[blockIdx.x] = c;
    }
}

// Exclusive scan of the per-block counts; offsets[nb] = total.
__global__ __launch_bounds__(kScanThreads) void k_psf_scan(const int* __restrict__ counts,
                                                           int nb, int* __restrict__ offsets)
{
    __shared__ int s[kScanThreads];
    const int per = (nb + kScanThreads - 1) / kScanThreads;
    const int b0 = threadIdx.x * per;
    int sum = 0;
    for (int b = b0; b < b0 + per && b < nb; b++) sum += counts[b];
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int d = 1; d < kScanThreads; d <<= 1) {
        int v = (threadIdx.x >= d) ? s[threadIdx.x - d] : 0;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    int run = s[threadIdx.x] - sum;
    for (int b = b0; b < b0 + per && b < nb; b++) {
        offsets[b] = run;
        run += counts[b];
    }
    if (threadIdx.x == kScanThreads - 1) offsets[nb] = s[kScanThreads - 1];
}

// Per ray: phase at p0 (revolutions) and per-pixel phase steps
//   o   = [alpha_hi, alpha_lo, beta_hi, beta_lo, phi0_rev, flux, 0, 0]
// and the rotations between the row blocks of a wave's tile (16, 32, 48
// rows apart), with which the split-BF16 path derives blocks 1-3 from 0
//   rot = [cos, sin of 2 pi frac(16 k alpha) for k = 1, 2, 3, 0, 0].
__device__ __forceinline__ void prep_ray(const bt_rays& rv, int64_t i, double p0x, double p0y,
                                         double tref, double a1x, double a1y, double a2x,
                                         double a2y, int keep_a, int keep_b,
                                         float* __restrict__ o, float* __restrict__ rot)
{
    const double x = rv.x[i], y = rv.y[i], z = rv.z[i];
    const double vx = rv.vx[i], vy = rv.vy[i], vz = rv.vz[i];
    const double t = rv.t[i], w_ = rv.wavelength[i];
    // the reference's phase at p0 (rayVector.cpp:207-214), z_p = 0
    const double v2 = vx*vx + vy*vy + vz*vz;
    double ph = (p0x-x)*vx;
    ph += (p0y-y)*vy;
    ph += (0.0-z)*vz;
    ph /= v2;
    ph -= tref-t;
    ph *= kTwoPi / w_;
    double rev = ph / kTwoPi;
    rev -= floor(rev);
    // revolutions per pixel step along a1 and a2
    const double inv = 1.0 / (v2 * w_);
    const double al = (a1x*vx + a1y*vy) * inv;
    const double be = (a2x*vx + a2y*vy) * inv;
    const float ah = hi_bits((float)al, keep_a);
    const float bh = hi_bits((float)be, keep_b);
    o[0] = ah;
    o[1] = (float)(al - (double)ah);
    o[2] = bh;
    o[3] = (float)(be - (double)bh);
    o[4] = (float)rev;
    o[5] = (float)rv.flux[i];
    o[6] = 0.f;
    o[7] = 0.f;
    for (int k = 1; k <= 3; k++) {
        double d = 16.0 * k * al;
        d -= floor(d);
        rot[2*k - 2] = (float)cos(kTwoPi * d);
        rot[2*k - 1] = (float)sin(kTwoPi * d);
    }
    rot[6] = 0.f;
    rot[7] = 0.f;
}

// The parameters of every valid ray, compacted in ray order.
__global__ __launch_bounds__(kBlock) void k_psf_prep(bt_rays rv, PsfGeom g, int ignore_vig,
                                                      const int* __restrict__ offsets,
                                                      float* __restrict__ params,
                                                      float* __restrict__ rot)
{
    __shared__ int s[kBlock / 64];
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool v = i < rv.n && psf_valid(rv, i, ignore_vig);
    const uint64_t m = __ballot(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) s[w] = __popcll(m);
    __syncthreads();
    if (!v) return;
    int base = offsets[blockIdx.x];
    for (int k = 0; k < w; k++) base += s[k];
    const int rank = base + __popcll(m & ((1ull << lane) - 1ull));
    prep_ray(rv, i, g.p0x, g.p0y, g.t, g.a1x, g.a1y, g.a2x, g.a2y, g.keep_a, g.keep_b,
             params + (int64_t)rank * kRayStride, rot + (int64_t)rank * kRotStride);
}

// sin(2 pi theta) with theta in revolutions (v_sin_f32).
__device__ __forceinline__ float sin_rev(float th) { return __builtin_amdgcn_sinf(th); }
__device__ __forceinline__ float cos_rev(float th) { return __builtin_amdgcn_cosf(th); }

__device__ __forceinline__ float step_phase(float c, float hi, float lo, float add)
{
    // fract(c*hi) is exact: c*hi is exact by construction of hi
    return __builtin_amdgcn_fractf(c * hi) + __builtin_fmaf(c, lo, add);
}

// One workgroup's 128 x 128 complex tile of the sum over rays [r0, r1) of
// the prepared parameters, written as FP32 (re, im) into out[i][j][2] (the
// padded nxp x nyp image).  Every lane of the workgroup must call it.
template <int kPrec>
__device__ __forceinline__ void psf_tile(const float* __restrict__ params,
                                         const float* __restrict__ rot, int r0, int r1,
                                         int tile_i0, int tile_j0, int nx, int ny, int nyp,
                                         float* __restrict__ out)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int i0 = tile_i0 + (wave & 1) * 64;
    const int j0 = tile_j0 + (wave >> 1) * 64;
    const int half = lane >> 5;       // k half of the operand
    const int l32 = lane & 31;
    v16f acc[4][2];
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < 2; b++)
            for (int e = 0; e < 16; e++) acc[a][b][e] = 0.f;

    // per-lane pixel offsets of the A rows (i) and B columns (j)
    float ci[4], cj[2];
    const int prow = l32 & 1;                          // re / im row
    for (int rb = 0; rb < 4; rb++) ci[rb] = (float)(i0 + rb*16 + (l32 >> 1) - nx/2);
    for (int cb = 0; cb < 2; cb++) cj[cb] = (float)(j0 + cb*32 + l32 - ny/2);

    if (kPrec == BT_PSF_FP32) {
        // 32x32x2: lane supplies A'[l32][half], B'[half][l32]; one ray per MFMA
        //   (prow, half) = (0,0): Re U = cos = sin(th + 1/4)
        //                  (0,1): -Im U = -sin = sin(th + 1/2)
        //                  (1,0): Im U = sin(th);  (1,1): Re U = sin(th + 1/4)
        // (th in revolutions), so every A value is one sine with a per-lane
        // offset; the ray's flux scales its B values (Re V = cos, Im V = sin).
        const float oa = (prow == half) ? 0.25f : (half == 1 ? 0.5f : 0.f);
        const float ob = (half == 0) ? 0.25f : 0.f;
        // rays in groups of 4 whose parameters (wave-uniform) are fetched one
        // group ahead of their use
        constexpr int G = 4;
        float cur[G][6], nxt[G][6];
        auto fetch = [&](float (&dst)[G][6], int rbase) {
            for (int q = 0; q < G; q++) {
                const int rr = rbase + q;
                const bool live = rr < r1;
                const float* p = params + (int64_t)(live ? rr : r0) * kRayStride;
                for (int e = 0; e < 5; e++) dst[q][e] = p[e];
                dst[q][5] = live ? p[5] : 0.f;
            }
        };
        if (r0 < r1) fetch(cur, r0);
        for (int r = r0; r < r1; r += G) {
            if (r + G < r1) fetch(nxt, r + G);
            for (int q = 0; q < G; q++) {
                const float ah = cur[q][0], al = cur[q][1], bh = cur[q][2], bl = cur[q][3];
                const float pa = cur[q][4] + oa, f = cur[q][5];
                float av[4], bv[2];
                for (int rb = 0; rb < 4; rb++)
                    av[rb] = sin_rev(step_phase(ci[rb], ah, al, pa));
                for (int cb = 0; cb < 2; cb++)
                    bv[cb] = f * sin_rev(step_phase(cj[cb], bh, bl, ob));
                for (int rb = 0; rb < 4; rb++)
                    for (int cb = 0; cb < 2; cb++)
                        acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[rb], bv[cb],
                                                                          acc[rb][cb], 0, 0, 0);
            }
            for (int q = 0; q < G; q++)
                for (int e = 0; e < 6; e++) cur[q][e] = nxt[q][e];
        }
    } else {
        // 32x32x16 bf16: lane supplies A'[l32][8*half + e], e = 0..7, i.e. rays
        // 4*half + e/2, part e%2; same for B'.
        for (int r = r0; r < r1; r += kRayGroup) {
            v8bf ahi[4], alo[4], bhi[2], blo[2];
            for (int q = 0; q < 4; q++) {
                const int rr = r + 4*half + q;
                const bool live = rr < r1;
                const float* p = params + (int64_t)(live ? rr : r0) * kRayStride;
                const float ah = p[0], al = p[1], bh = p[2], bl = p[3], ph0 = p[4];
                const float f = live ? p[5] : 0.f;
                // row block 0 evaluated, blocks 1-3 rotated from it
                const float th = step_phase(ci[0], ah, al, ph0);
                const float fs0 = f * sin_rev(th), fc0 = f * cos_rev(th);
                for (int rb = 0; rb < 4; rb++) {
                    float s = fs0, c = fc0;
                    if (rb > 0) {
                        const float* pr = rot + (int64_t)(live ? rr : r0) * kRotStride;
                        const float cr = pr[2*rb - 2], sr = pr[2*rb - 1];
                        s = __builtin_fmaf(fs0, cr, fc0 * sr);
                        c = __builtin_fmaf(fc0, cr, -(fs0 * sr));
                    }
                    const float v0 = prow ? s : c;        // k part 0
                    const float v1 = prow ? c : -s;       // k part 1
                    const __bf16 h0 = (__bf16)v0, h1 = (__bf16)v1;
                    ahi[rb][2*q] = h0;
                    ahi[rb][2*q+1] = h1;
                    alo[rb][2*q] = (__bf16)(v0 - (float)h0);
                    alo[rb][2*q+1] = (__bf16)(v1 - (float)h1);
                }
                for (int cb = 0; cb < 2; cb++) {
                    const float th = step_phase(cj[cb], bh, bl, 0.f);
                    const float c = cos_rev(th), s = sin_rev(th);
                    const __bf16 hc = (__bf16)c, hs = (__bf16)s;
                    bhi[cb][2*q] = hc;
                    bhi[cb][2*q+1] = hs;
                    blo[cb][2*q] = (__bf16)(c - (float)hc);
                    blo[cb][2*q+1] = (__bf16)(s - (float)hs);
                }
            }
            for (int rb = 0; rb < 4; rb++)
                for (int cb = 0; cb < 2; cb++) {
                    v16f a = acc[rb][cb];
                    a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(alo[rb], bhi[cb], a, 0, 0, 0);
                    a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[rb], blo[cb], a, 0, 0, 0);
                    a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ahi[rb], bhi[cb], a, 0, 0, 0);
                    acc[rb][cb] = a;
                }
        }
    }

    // D[row][col]: row = 8*(e/4) + 4*half + e%4, col = l32 (32x32 MFMA layout)
    // out[i][j][2] over the padded nxp x nyp image
    for (int rb = 0; rb < 4; rb++)
        for (int cb = 0; cb < 2; cb++)
            for (int e = 0; e < 16; e++) {
                const int row = 8*(e >> 2) + 4*half + (e & 3);
                const int i = i0 + rb*16 + (row >> 1);
                const int j = j0 + cb*32 + l32;
                out[((int64_t)i * nyp + j) * 2 + (row & 1)] = acc[rb][cb][e];
            }
}


template <int kPrec>
__global__ __launch_bounds__(256, 2) void k_psf_gemm(
    const float* __restrict__ params, const float* __restrict__ rot,
    const int* __restrict__ total_ptr, int nx, int ny, int tiles_i, int ksplit, int nxp,
    int nyp, float* __restrict__ partial)
{
    const int tile = blockIdx.x;
    const int ks = blockIdx.y;
    const int total = *total_ptr;
    int per = (total + ksplit - 1) / ksplit;
    per = ((per + kRayGroup - 1) / kRayGroup) * kRayGroup;
    const int r0 = ks * per;
    const int r1 = min(total, r0 + per);
    psf_tile<kPrec>(params, rot, r0, r1, (tile % tiles_i) * kTile, (tile / tiles_i) * kTile, nx, ny,
                    nyp, partial + (int64_t)ks * nxp * nyp * 2);
}

// amp[j][i] = sum_ks partial[ks][i][j] in FP64, in a fixed order: a block
// covers 32 consecutive pixels (j fastest, coalesced float2 reads) with 8
// lanes per pixel, lane g summing splits g, g+8, ...; the 8 sums are then
// added in g order.
constexpr int kRedPix = 32;
constexpr int kRedLanes = kBlock / kRedPix;

__device__ __forceinline__ void reduce_splits(const float2* __restrict__ p, int64_t stride,
                                              int k0, int k1, bool live, double* __restrict__ out)
{
    __shared__ double s_re[kRedLanes][kRedPix], s_im[kRedLanes][kRedPix];
    const int pj = threadIdx.x % kRedPix, g = threadIdx.x / kRedPix;
    double re = 0.0, im = 0.0;
    if (live)
        for (int k = k0 + g; k < k1; k += kRedLanes) {
            const float2 v = p[k * stride];
            re += (double)v.x;
            im += (double)v.y;
        }
    s_re[g][pj] = re;
    s_im[g][pj] = im;
    __syncthreads();
    if (g == 0 && live) {
        double r = 0.0, m = 0.0;
        for (int q = 0; q < kRedLanes; q++) { r += s_re[q][pj]; m += s_im[q][pj]; }
        out[0] = r;
        out[1] = m;
    }
}

__global__ __launch_bounds__(kBlock) void k_psf_reduce(const float* __restrict__ partial,
                                                       int ksplit, int nx, int ny, int nxp,
                                                       int nyp, double* __restrict__ amp)
{
    const int64_t q = (int64_t)blockIdx.x * kRedPix + threadIdx.x % kRedPix;
    const bool live = q < (int64_t)nx * ny;
    const int i = live ? (int)(q / ny) : 0, j = live ? (int)(q % ny) : 0;
    const float2* p = reinterpret_cast<const float2*>(partial) + ((int64_t)i * nyp + j);
    reduce_splits(p, (int64_t)nxp * nyp, 0, ksplit, live, amp + 2 * ((int64_t)j * nx + i));
}

// ---------------------------------------------------------------------------
// Batched PSFs: many (field, wavelength) segments of one ray batch, each with
// its own lattice, reference point and time, summed in one persistent GEMM
// launch.  Capture-safe: no host synchronisation, allocation or memset.
// ---------------------------------------------------------------------------
struct SegGeom {
    double p0x, p0y, t, pad;
};

// Per segment: reference point (the chief ray's position, or the mean of the
// unvignetted rays' positions, analysis.py:289-295) and time (the t of
// `time_ray`, analysis.py:301).  One block per segment; the mean is summed in
// a fixed order.
__global__ __launch_bounds__(kBlock) void k_psf_seg_geom(bt_rays rv,
                                                          const bt_psf_segment* __restrict__ segs,
                                                          SegGeom* __restrict__ geom)
{
    __shared__ double sx[kBlock], sy[kBlock];
    __shared__ int64_t sc[kBlock];
    const bt_psf_segment sg = segs[blockIdx.x];
    const int64_t b = sg.begin < 0 ? 0 : (sg.begin > rv.n ? rv.n : sg.begin);
    const int64_t e = sg.end < b ? b : (sg.end > rv.n ? rv.n : sg.end);
    double px, py;
    if (sg.chief >= 0 && sg.chief < rv.n) {
        px = rv.x[sg.chief];
        py = rv.y[sg.chief];
    } else {
        double ax = 0.0, ay = 0.0;
        int64_t c = 0;
        for (int64_t i = b + threadIdx.x; i < e; i += kBlock)
            if (!rv.vignetted[i]) { ax += rv.x[i]; ay += rv.y[i]; c++; }
        sx[threadIdx.x] = ax;
        sy[threadIdx.x] = ay;
        sc[threadIdx.x] = c;
        __syncthreads();
        for (int d = kBlock / 2; d > 0; d >>= 1) {
            if ((int)threadIdx.x < d) {
                sx[threadIdx.x] += sx[threadIdx.x + d];
                sy[threadIdx.x] += sy[threadIdx.x + d];
                sc[threadIdx.x] += sc[threadIdx.x + d];
            }
            __syncthreads();
        }
        px = sx[0] / (double)sc[0];
        py = sy[0] / (double)sc[0];
    }
    if (threadIdx.x == 0) {
        const int64_t tr = (sg.time_ray >= 0 && sg.time_ray < rv.n) ? sg.time_ray : b;
        geom[blockIdx.x] = SegGeom{px, py, tr < rv.n ? rv.t[tr] : 0.0, 0.0};
    }
}

// Segment of ray i: the last segment whose begin <= i, if i < its end; -1
// when no segment holds the ray.  Segments are sorted by begin.
__device__ __forceinline__ int find_segment(const bt_psf_segment* __restrict__ segs, int nseg,
                                            int64_t i)
{
    int lo = 0, hi = nseg - 1, s = -1;
    while (lo <= hi) {
        const int mid = (lo + hi) >> 1;
        if (segs[mid].begin <= i) { s = mid; lo = mid + 1; } else { hi = mid - 1; }
    }
    return (s >= 0 && i < segs[s].end) ? s : -1;
}

// As k_psf_prep, each ray with its own segment's geometry.
__global__ __launch_bounds__(kBlock) void k_psf_prep_batch(
    bt_rays rv, const bt_psf_segment* __restrict__ segs, int nseg,
    const SegGeom* __restrict__ geom, int keep_a, int keep_b, int ignore_vig,
    const int* __restrict__ offsets, float* __restrict__ params, float* __restrict__ rot)
{
    __shared__ int s[kBlock / 64];
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool v = i < rv.n && psf_valid(rv, i, ignore_vig);
    const uint64_t m = __ballot(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) s[w] = __popcll(m);
    __syncthreads();
    if (!v) return;
    int base = offsets[blockIdx.x];
    for (int k = 0; k < w; k++) base += s[k];
    const int rank = base + __popcll(m & ((1ull << lane) - 1ull));
    float* o = params + (int64_t)rank * kRayStride;
    float* orot = rot + (int64_t)rank * kRotStride;
    const int sg = find_segment(segs, nseg, i);
    if (sg < 0) {                       // outside every segment: never summed
        for (int k = 0; k < kRayStride; k++) o[k] = 0.f;
        for (int k = 0; k < kRotStride; k++) orot[k] = 0.f;
        return;
    }
    const SegGeom g = geom[sg];
    const bt_psf_segment& d = segs[sg];
    prep_ray(rv, i, g.p0x, g.p0y, g.t, d.a1[0], d.a1[1], d.a2[0], d.a2[1], keep_a, keep_b, o,
             orot);
}

// Number of valid rays before ray index b (b <= n), from the block offsets.
__device__ __forceinline__ int valid_before(const bt_rays& rv, const int* __restrict__ offsets,
                                            int64_t b, int ignore_vig)
{
    const int64_t blk = b / kBlock;
    int c = offsets[blk];
    for (int64_t i = blk * kBlock; i < b; i++) c += psf_valid(rv, i, ignore_vig) ? 1 : 0;
    return c;
}

struct SegRange {
    int begin;      // first compacted ray of the segment
    int count;      // compacted rays of the segment
    int items;      // work items (chunks of `chunk` rays)
    int first;      // index of the segment's first work item
};

__global__ __launch_bounds__(kBlock) void k_psf_seg_ranges(
    bt_rays rv, const bt_psf_segment* __restrict__ segs, int nseg, int ignore_vig,
    const int* __restrict__ offsets, int chunk, SegRange* __restrict__ ranges)
{
    const int sg = blockIdx.x * kBlock + threadIdx.x;
    if (sg >= nseg) return;
    const bt_psf_segment d = segs[sg];
    const int64_t b = d.begin < 0 ? 0 : (d.begin > rv.n ? rv.n : d.begin);
    const int64_t e = d.end < b ? b : (d.end > rv.n ? rv.n : d.end);
    const int cb = valid_before(rv, offsets, b, ignore_vig);
    const int ce = valid_before(rv, offsets, e, ignore_vig);
    SegRange r;
    r.begin = cb;
    r.count = ce - cb;
    r.items = (r.count + chunk - 1) / chunk;
    r.first = 0;
    ranges[sg] = r;
}

struct WorkItem {
    int seg, r0, r1, pad;
};

// Exclusive scan of the segments' item counts, then the item list
// (segment-major, chunks in ray order); items[cap..] are never written and
// *nitems is clamped to cap.
__global__ __launch_bounds__(kScanThreads) void k_psf_items(SegRange* __restrict__ ranges,
                                                            int nseg, int chunk, int cap,
                                                            WorkItem* __restrict__ items,
                                                            int* __restrict__ nitems)
{
    __shared__ int s[kScanThreads];
    const int per = (nseg + kScanThreads - 1) / kScanThreads;
    const int s0 = threadIdx.x * per;
    int sum = 0;
    for (int q = s0; q < s0 + per && q < nseg; q++) sum += ranges[q].items;
    s[threadIdx.x] = sum;
    __syncthreads();
    for (int d = 1; d < kScanThreads; d <<= 1) {
        int v = (threadIdx.x >= d) ? s[threadIdx.x - d] : 0;
        __syncthreads();
        s[threadIdx.x] += v;
        __syncthreads();
    }
    int run = s[threadIdx.x] - sum;
    for (int q = s0; q < s0 + per && q < nseg; q++) {
        SegRange r = ranges[q];
        r.first = run;
        ranges[q] = r;
        for (int c = 0; c < r.items; c++) {
            const int it = run + c;
            if (it >= cap) break;
            const int r0 = r.begin + c * chunk;
            const int r1 = min(r.begin + r.count, r0 + chunk);
            items[it] = WorkItem{q, r0, r1, 0};
        }
        run += r.items;
    }
    if (threadIdx.x == kScanThreads - 1) {
        const int total = s[kScanThreads - 1];
        *nitems = total < cap ? total : cap;
    }
}

// Persistent GEMM: workgroup w takes (item, tile) pairs w, w + grid, ... in a
// fixed assignment, so the result does not depend on scheduling.
template <int kPrec>
__global__ __launch_bounds__(256, 2) void k_psf_gemm_batch(
    const float* __restrict__ params, const float* __restrict__ rot,
    const WorkItem* __restrict__ items,
    const int* __restrict__ nitems_ptr, int tiles_i, int tiles, int nx, int ny, int nxp,
    int nyp, float* __restrict__ partial)
{
    const int64_t nwork = (int64_t)(*nitems_ptr) * tiles;
    for (int64_t wk = blockIdx.x; wk < nwork; wk += gridDim.x) {
        const int item = (int)(wk / tiles);
        const int tile = (int)(wk % tiles);
        const WorkItem it = items[item];
        psf_tile<kPrec>(params, rot, it.r0, it.r1, (tile % tiles_i) * kTile, (tile / tiles_i) * kTile,
                        nx, ny, nyp, partial + (int64_t)item * nxp * nyp * 2);
    }
}

// amp[s][j][i] = sum over the segment's items of partial[item][i][j], FP64,
// in the fixed order of reduce_splits.
__global__ __launch_bounds__(kBlock) void k_psf_reduce_batch(
    const float* __restrict__ partial, const SegRange* __restrict__ ranges, int nseg, int nx,
    int ny, int nxp, int nyp, int cap, double* __restrict__ amp)
{
    const int64_t npix = (int64_t)nx * ny;
    const int64_t blocks_per_seg = (npix + kRedPix - 1) / kRedPix;
    const int sg = (int)(blockIdx.x / blocks_per_seg);
    const int64_t q = (int64_t)(blockIdx.x % blocks_per_seg) * kRedPix + threadIdx.x % kRedPix;
    const bool live = q < npix;
    const int i = live ? (int)(q / ny) : 0, j = live ? (int)(q % ny) : 0;
    const SegRange r = ranges[sg];
    const int last = min(r.first + r.items, cap);
    const float2* p = reinterpret_cast<const float2*>(partial) + ((int64_t)i * nyp + j);
    reduce_splits(p, (int64_t)nxp * nyp, r.first, last, live,
                  amp + 2 * ((int64_t)sg * npix + (int64_t)j * nx + i));
}

struct GemmLayout {
    int64_t nblocks;      // ray blocks of kBlock
    int tiles_i, tiles_j, nxp, nyp, ksplit;
    int64_t off_params, off_counts, off_offsets, off_partial, bytes;
};

int bits_for(int c)
{
    int b = 0;
    while ((1 << b) <= c) b++;
    return b;
}

GemmLayout gemm_layout(int64_t n, int nx, int ny)
{
    GemmLayout L;
    L.nblocks = (n + kBlock - 1) / kBlock;
    if (L.nblocks < 1) L.nblocks = 1;
    L.tiles_i = (nx + kTile - 1) / kTile;
    L.tiles_j = (ny + kTile - 1) / kTile;
    L.nxp = L.tiles_i * kTile;
    L.nyp = L.tiles_j * kTile;
    const int tiles = L.tiles_i * L.tiles_j;
    // about 512 workgroups (2048 waves: two per SIMD), at least 64 rays each
    int ks = (512 + tiles - 1) / tiles;
    int64_t by_rays = (n + 63) / 64;
    if (by_rays < 1) by_rays = 1;
    if (ks > by_rays) ks = (int)by_rays;
    if (ks > 4096) ks = 4096;
    if (ks < 1) ks = 1;
    L.ksplit = ks;
    auto align = [](int64_t v) { return (v + 255) & ~(int64_t)255; };
    int64_t off = 0;
    L.off_params = off;
    off = align(off + (n + kRayGroup) * (kRayStride + kRotStride) * (int64_t)sizeof(float));
    L.off_counts = off;  off = align(off + L.nblocks * (int64_t)sizeof(int));
    L.off_offsets = off; off = align(off + (L.nblocks + 1) * (int64_t)sizeof(int));
    L.off_partial = off; off = align(off + (int64_t)ks * L.nxp * L.nyp * 2 * (int64_t)sizeof(float));
    L.bytes = off;
    return L;
}

}  // namespace bt

using namespace bt;

extern "C" int64_t bt_huygens_workspace_bytes(int64_t n, int nx, int ny)
{
    if (n < 0 || nx <= 0 || ny <= 0) return 0;
    const int64_t d = bt_huygens_direct_workspace_bytes(n, nx, ny);
    const int64_t g = gemm_layout(n, nx, ny).bytes;
    return d > g ? d : g;
}

extern "C" int bt_huygens_gemm(bt_rays rays, const double* p0, const double* a1,
                               const double* a2, int nx, int ny, double t,
                               int ignore_vignetted, int precision, double* amp_out,
                               void* workspace, int64_t workspace_bytes, void* stream)
{
    if (!rays_ok(rays) || !p0 || !a1 || !a2 || !amp_out || !workspace)
        return fail(BT_E_NULL, "bt_huygens_gemm: null argument");
    if (nx <= 0 || ny <= 0 || nx > 8192 || ny > 8192)
        return fail(BT_E_SIZE, "bt_huygens_gemm: image size must be 1..8192 per side");
    if (rays.n > (int64_t)1 << 30)
        return fail(BT_E_SIZE, "bt_huygens_gemm: more than 2^30 rays");
    if (precision != BT_PSF_FP32 && precision != BT_PSF_BF16X3)
        return fail(BT_E_KIND, "bt_huygens_gemm: unknown precision");
    const GemmLayout L = gemm_layout(rays.n, nx, ny);
    if (workspace_bytes < L.bytes) return fail(BT_E_SIZE, "bt_huygens_gemm: workspace too small");
    char* ws = static_cast<char*>(workspace);
    float* params = reinterpret_cast<float*>(ws + L.off_params);
    float* rot = params + (rays.n + kRayGroup) * kRayStride;
    int* counts = reinterpret_cast<int*>(ws + L.off_counts);
    int* offsets = reinterpret_cast<int*>(ws + L.off_offsets);
    float* partial = reinterpret_cast<float*>(ws + L.off_partial);
    hipStream_t s = as_stream(stream);

    PsfGeom g;
    g.p0x = p0[0]; g.p0y = p0[1];
    g.a1x = a1[0]; g.a1y = a1[1];
    g.a2x = a2[0]; g.a2y = a2[1];
    g.t = t;
    g.nx = nx; g.ny = ny;
    // |ci| <= max(nx/2, nxp - nx/2): keep 24 - bits(|ci|) bits so ci*hi is exact
    const int cmax_i = (L.nxp - nx / 2 > nx / 2) ? L.nxp - nx / 2 : nx / 2;
    const int cmax_j = (L.nyp - ny / 2 > ny / 2) ? L.nyp - ny / 2 : ny / 2;
    g.keep_a = 24 - bits_for(cmax_i);
    g.keep_b = 24 - bits_for(cmax_j);

    const unsigned nb = (unsigned)L.nblocks;
    if (rays.n > 0) {
        hipLaunchKernelGGL(k_psf_count, dim3(nb), dim3(kBlock), 0, s, rays, ignore_vignetted, counts);
        int rc = check_launch("bt_huygens_gemm(count)");
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_psf_scan, dim3(1), dim3(kScanThreads), 0, s, counts,
                       rays.n > 0 ? (int)nb : 0, offsets);
    int rc = check_launch("bt_huygens_gemm(scan)");
    if (rc) return rc;
    if (rays.n > 0) {
        hipLaunchKernelGGL(k_psf_prep, dim3(nb), dim3(kBlock), 0, s, rays, g, ignore_vignetted,
                           offsets, params, rot);
        rc = check_launch("bt_huygens_gemm(prep)");
        if (rc) return rc;
    }
    const int* total = offsets + (rays.n > 0 ? nb : 0);
    dim3 grid(L.tiles_i * L.tiles_j, L.ksplit);
    if (precision == BT_PSF_FP32)
        hipLaunchKernelGGL(k_psf_gemm<BT_PSF_FP32>, grid, dim3(256), 0, s, params, rot, total, nx, ny,
                           L.tiles_i, L.ksplit, L.nxp, L.nyp, partial);
    else
        hipLaunchKernelGGL(k_psf_gemm<BT_PSF_BF16X3>, grid, dim3(256), 0, s, params, rot, total, nx, ny,
                           L.tiles_i, L.ksplit, L.nxp, L.nyp, partial);
    rc = check_launch("bt_huygens_gemm(mfma)");
    if (rc) return rc;
    hipLaunchKernelGGL(k_psf_reduce, dim3((unsigned)(((int64_t)nx * ny + kRedPix - 1) / kRedPix)),
                       dim3(kBlock), 0, s, partial, L.ksplit, nx, ny, L.nxp, L.nyp, amp_out);
    return check_launch("bt_huygens_gemm(reduce)");
}

namespace bt {

struct BatchLayout {
    int64_t nblocks, cap;
    int tiles_i, tiles_j, nxp, nyp;
    int64_t off_geom, off_counts, off_offsets, off_ranges, off_items, off_nitems, off_params,
        off_partial, bytes;
};

BatchLayout batch_layout(int64_t n, int nseg, int nx, int ny, int chunk)
{
    BatchLayout L;
    L.nblocks = (n + kBlock - 1) / kBlock;
    if (L.nblocks < 1) L.nblocks = 1;
    L.cap = (n + chunk - 1) / chunk + nseg;
    L.tiles_i = (nx + kTile - 1) / kTile;
    L.tiles_j = (ny + kTile - 1) / kTile;
    L.nxp = L.tiles_i * kTile;
    L.nyp = L.tiles_j * kTile;
    auto align = [](int64_t v) { return (v + 255) & ~(int64_t)255; };
    int64_t off = 0;
    L.off_geom = off;    off = align(off + nseg * (int64_t)sizeof(SegGeom));
    L.off_counts = off;  off = align(off + L.nblocks * (int64_t)sizeof(int));
    L.off_offsets = off; off = align(off + (L.nblocks + 1) * (int64_t)sizeof(int));
    L.off_ranges = off;  off = align(off + nseg * (int64_t)sizeof(SegRange));
    L.off_items = off;   off = align(off + L.cap * (int64_t)sizeof(WorkItem));
    L.off_nitems = off;  off = align(off + (int64_t)sizeof(int));
    L.off_params = off;
    off = align(off + (n + kRayGroup) * (kRayStride + kRotStride) * (int64_t)sizeof(float));
    L.off_partial = off; off = align(off + L.cap * L.nxp * L.nyp * 2 * (int64_t)sizeof(float));
    L.bytes = off;
    return L;
}

bool batch_args_ok(int64_t n, int nseg, int nx, int ny, int chunk)
{
    return n >= 0 && n <= ((int64_t)1 << 30) && nseg >= 1 && nseg <= (1 << 20) && nx >= 1 &&
           ny >= 1 && nx <= 8192 && ny <= 8192 && chunk >= kRayGroup && chunk <= (1 << 24) &&
           chunk % kRayGroup == 0 && (int64_t)nx * ny * nseg <= ((int64_t)1 << 31);
}

}  // namespace bt

extern "C" int64_t bt_huygens_batch_workspace_bytes(int64_t n, int nseg, int nx, int ny,
                                                    int chunk)
{
    if (!batch_args_ok(n, nseg, nx, ny, chunk)) return 0;
    return batch_layout(n, nseg, nx, ny, chunk).bytes;
}

extern "C" int bt_huygens_batch(bt_rays rays, const bt_psf_segment* segments, int nseg, int nx,
                                int ny, int ignore_vignetted, int precision, int chunk,
                                int nwg, double* amp_out, void* workspace,
                                int64_t workspace_bytes, void* stream)
{
    if (!rays_ok(rays) || !segments || !amp_out || !workspace)
        return fail(BT_E_NULL, "bt_huygens_batch: null argument");
    if (!batch_args_ok(rays.n, nseg, nx, ny, chunk))
        return fail(BT_E_SIZE, "bt_huygens_batch: sizes out of range (rays <= 2^30, "
                               "1 <= segments <= 2^20, image 1..8192 per side, chunk a "
                               "multiple of 8)");
    if (precision != BT_PSF_FP32 && precision != BT_PSF_BF16X3)
        return fail(BT_E_KIND, "bt_huygens_batch: unknown precision");
    if (nwg < 0 || nwg > (1 << 20)) return fail(BT_E_SIZE, "bt_huygens_batch: bad workgroup count");
    const BatchLayout L = batch_layout(rays.n, nseg, nx, ny, chunk);
    if (workspace_bytes < L.bytes) return fail(BT_E_SIZE, "bt_huygens_batch: workspace too small");
    char* ws = static_cast<char*>(workspace);
    SegGeom* geom = reinterpret_cast<SegGeom*>(ws + L.off_geom);
    int* counts = reinterpret_cast<int*>(ws + L.off_counts);
    int* offsets = reinterpret_cast<int*>(ws + L.off_offsets);
    SegRange* ranges = reinterpret_cast<SegRange*>(ws + L.off_ranges);
    WorkItem* items = reinterpret_cast<WorkItem*>(ws + L.off_items);
    int* nitems = reinterpret_cast<int*>(ws + L.off_nitems);
    float* params = reinterpret_cast<float*>(ws + L.off_params);
    float* rot = params + (rays.n + kRayGroup) * kRayStride;
    float* partial = reinterpret_cast<float*>(ws + L.off_partial);
    hipStream_t s = as_stream(stream);

    const int cmax_i = (L.nxp - nx / 2 > nx / 2) ? L.nxp - nx / 2 : nx / 2;
    const int cmax_j = (L.nyp - ny / 2 > ny / 2) ? L.nyp - ny / 2 : ny / 2;
    const int keep_a = 24 - bits_for(cmax_i);
    const int keep_b = 24 - bits_for(cmax_j);
    const unsigned nb = (unsigned)L.nblocks;
    const int nbk = rays.n > 0 ? (int)nb : 0;

    hipLaunchKernelGGL(k_psf_seg_geom, dim3(nseg), dim3(kBlock), 0, s, rays, segments, geom);
    int rc = check_launch("bt_huygens_batch(geometry)");
    if (rc) return rc;
    if (rays.n > 0) {
        hipLaunchKernelGGL(k_psf_count, dim3(nb), dim3(kBlock), 0, s, rays, ignore_vignetted,
                           counts);
        rc = check_launch("bt_huygens_batch(count)");
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_psf_scan, dim3(1), dim3(kScanThreads), 0, s, counts, nbk, offsets);
    rc = check_launch("bt_huygens_batch(scan)");
    if (rc) return rc;
    if (rays.n > 0) {
        hipLaunchKernelGGL(k_psf_prep_batch, dim3(nb), dim3(kBlock), 0, s, rays, segments, nseg,
                           geom, keep_a, keep_b, ignore_vignetted, offsets, params, rot);
        rc = check_launch("bt_huygens_batch(prep)");
        if (rc) return rc;
    }
    hipLaunchKernelGGL(k_psf_seg_ranges, dim3((nseg + kBlock - 1) / kBlock), dim3(kBlock), 0, s,
                       rays, segments, nseg, ignore_vignetted, offsets, chunk, ranges);
    rc = check_launch("bt_huygens_batch(ranges)");
    if (rc) return rc;
    hipLaunchKernelGGL(k_psf_items, dim3(1), dim3(kScanThreads), 0, s, ranges, nseg, chunk,
                       (int)L.cap, items, nitems);
    rc = check_launch("bt_huygens_batch(items)");
    if (rc) return rc;
    const int grid = nwg > 0 ? nwg : 512;
    const int tiles = L.tiles_i * L.tiles_j;
    if (precision == BT_PSF_FP32)
        hipLaunchKernelGGL(k_psf_gemm_batch<BT_PSF_FP32>, dim3(grid), dim3(256), 0, s, params,
                           rot, items, nitems, L.tiles_i, tiles, nx, ny, L.nxp, L.nyp, partial);
    else
        hipLaunchKernelGGL(k_psf_gemm_batch<BT_PSF_BF16X3>, dim3(grid), dim3(256), 0, s, params,
                           rot, items, nitems, L.tiles_i, tiles, nx, ny, L.nxp, L.nyp, partial);
    rc = check_launch("bt_huygens_batch(mfma)");
    if (rc) return rc;
    const int64_t red_blocks = (((int64_t)nx * ny + kRedPix - 1) / kRedPix) * nseg;
    hipLaunchKernelGGL(k_psf_reduce_batch, dim3((unsigned)red_blocks), dim3(kBlock), 0, s, partial,
                       ranges, nseg, nx, ny, L.nxp, L.nyp, (int)L.cap, amp_out);
    return check_launch("bt_huygens_batch(reduce)");
}
