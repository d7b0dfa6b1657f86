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
// layout.  A workgroup owns a 128 x 128 complex output tile and a contiguous
// range of rays: FP32 as two waves of 64 (i) x 128 (j) (sixteen 32 x 32
// accumulator blocks each), split-BF16 as four waves of 64 x 64 (eight).
// On gfx950 the FP32 MFMA does not co-execute with VALU work, so the FP32
// path spends as few VALU cycles per MFMA as it can: per ray and wave it
// evaluates one sine/cosine pair for the rows and one for the columns and
// derives the other blocks by rotations precomputed per ray in FP64.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <mutex>
#include <string>

#include "bt_device.h"
#include "bt_common.h"

extern "C" int64_t bt_huygens_direct_workspace_bytes(int64_t n, int nx, int ny);

namespace bt {

typedef float v16f __attribute__((ext_vector_type(16)));
typedef unsigned int v4u __attribute__((ext_vector_type(4)));
typedef unsigned int v2u __attribute__((ext_vector_type(2)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef __bf16 v2bf __attribute__((ext_vector_type(2)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

constexpr double kTwoPi = 2 * 3.14159265358979323846;
constexpr int kTile = 128;          // workgroup tile (complex) in i and j
constexpr int kRayStride = 8;       // floats per prepared ray
constexpr int kRotStride = 16;      // floats per ray of row/column-block rotations
constexpr int kRayGroup = 8;        // rays per bf16 MFMA step; splits align to it
constexpr int kScanThreads = 1024;
// buffer resource word 3 for raw 32-bit loads (gfx9 family)
constexpr int kBufferFlags = 0x00020000;

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
//   o   = [alpha_hi, alpha_lo, beta_hi, beta_lo, phi0_rev, flux,
//          cos, sin of 2 pi frac(32 beta)]
// and the rotations between the row blocks of a wave's tile (16, 32, 48
// pixels apart) and between its column blocks (32, 64, 96 pixels apart),
// with which blocks 1-3 are derived from block 0
//   rot = [cos, sin of 2 pi frac(16 k alpha) for k = 1, 2, 3, 0, 0,
//          cos, sin of 2 pi frac(32 k beta) for k = 1, 2, 3, 0, 0].
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
    // one FP64 sincos per axis; the 2- and 3-block rotations by products
    // (relative error ~1e-16, far below the FP32 rounding of the table)
    auto blocks = [](double step, float* o) {
        double d = step - floor(step);
        double s1, c1;
        sincos(kTwoPi * d, &s1, &c1);
        const double c2 = c1 * c1 - s1 * s1, s2 = 2.0 * s1 * c1;
        const double c3 = c2 * c1 - s2 * s1, s3 = s2 * c1 + c2 * s1;
        o[0] = (float)c1; o[1] = (float)s1;
        o[2] = (float)c2; o[3] = (float)s2;
        o[4] = (float)c3; o[5] = (float)s3;
    };
    blocks(16.0 * al, rot);
    blocks(32.0 * be, rot + 8);
    rot[6] = 0.f;
    rot[7] = 0.f;
    rot[14] = 0.f;
    rot[15] = 0.f;
    o[6] = rot[8];                  // the one column rotation split-BF16 uses
    o[7] = rot[9];
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

// x = hi + lo for a pair (x0, x1): hi and lo as dwords of two bf16 (element
// 0 in the low half), each rounded to nearest even by v_cvt_pk_bf16_f32.
__device__ __forceinline__ void split_pair(float x0, float x1, uint32_t& hi, uint32_t& lo)
{
    hi = __builtin_bit_cast(uint32_t, __builtin_convertvector((v2f){x0, x1}, v2bf));
    const float r0 = x0 - __uint_as_float(hi << 16);
    const float r1 = x1 - __uint_as_float(hi & 0xffff0000u);
    lo = __builtin_bit_cast(uint32_t, __builtin_convertvector((v2f){r0, r1}, v2bf));
}

__device__ __forceinline__ float step_phase(float c, float hi, float lo, float add)
{
    // fract(c*hi) is exact: c*hi is exact by construction of hi
    return __builtin_amdgcn_fractf(c * hi) + __builtin_fmaf(c, lo, add);
}

// Threads per workgroup of the GEMM kernels: two waves (FP32) or four
// (split-BF16) over the 128 x 128 tile, and the waves per SIMD each is
// compiled for: the FP32 wave's 256 accumulators need the whole register
// file (one wave per SIMD), split-BF16 runs two.
// Column blocks of 32 per wave in split-BF16: 4 (64 x 128 wave tiles, two
// waves per workgroup, one per SIMD) or 2 (64 x 64, four waves, two per
// SIMD).  The wider tile spends the row operands' generation on twice as
// many MFMAs.
#ifndef BT_BF16_CB
#define BT_BF16_CB 4
#endif
constexpr int kBf16Cb = BT_BF16_CB;
template <int kPrec>
constexpr int gemm_cb() { return kPrec == BT_PSF_FP32 ? 4 : kBf16Cb; }
template <int kPrec>
constexpr int gemm_threads() { return gemm_cb<kPrec>() == 4 ? 128 : 256; }
template <int kPrec>
constexpr int gemm_waves_per_simd() { return gemm_cb<kPrec>() == 4 ? 1 : 2; }

// One workgroup's 128 x 128 complex tile of the sum over rays [r0, r1) of
// the prepared parameters, written as FP32 (re, im) into out[i][j][2] (the
// padded nxp x nyp image).  Every lane of the workgroup must call it.
template <int kPrec>
__device__ __forceinline__ void psf_tile(const float* __restrict__ params,
                                         const float* __restrict__ rot, int r0, int r1,
                                         int tile_i0, int tile_j0, int nx, int ny, int nyp,
                                         float* __restrict__ out)
{
    // column blocks of 32 per wave: 4 (64 x 128) or 2 (split-BF16 64 x 64)
    constexpr int CB = gemm_cb<kPrec>();
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int i0 = tile_i0 + (wave & 1) * 64;
    const int j0 = tile_j0 + (wave >> 1) * (32 * CB);
    const int half = lane >> 5;       // k half of the operand
    const int l32 = lane & 31;
    v16f acc[4][CB];
    for (int a = 0; a < 4; a++)
        for (int b = 0; b < CB; b++)
            for (int e = 0; e < 16; e++) acc[a][b][e] = 0.f;

    // per-lane pixel offsets of the A rows (i) and B columns (j)
    float ci[4], cj[CB];
    const int prow = l32 & 1;                          // re / im row
    for (int rb = 0; rb < 4; rb++) ci[rb] = (float)(i0 + rb*16 + (l32 >> 1) - nx/2);
    for (int cb = 0; cb < CB; cb++) cj[cb] = (float)(j0 + cb*32 + l32 - ny/2);

    if constexpr (kPrec == BT_PSF_FP32) {
        // 32x32x2: lane supplies A'[l32][half], B'[half][l32]; one ray per MFMA
        //   (prow, half) = (0,0): Re U = cos = sin(th + 1/4)
        //                  (0,1): -Im U = -sin = sin(th + 1/2)
        //                  (1,0): Im U = sin(th);  (1,1): Re U = sin(th + 1/4)
        // (th in revolutions), so every A value is one sine with a per-lane
        // offset, and so is every B value (Re V = cos, Im V = sin).  Block 0
        // is evaluated as s = sin, c = cos of its phase; block k is then
        // sin(th + k d) = s cos(k d) + c sin(k d) with the ray's rotations.
        // The flux scales the A values.
        const float oa = (prow == half) ? 0.25f : (half == 1 ? 0.5f : 0.f);
        const float ob = (half == 0) ? 0.25f : 0.f;
        // A ray's parameters (wave-uniform) as buffer loads into VGPRs, only
        // the 18 dwords used: [ah, al, bh, bl], [phi0, flux], row rotations
        // [c1, s1, c2, s2], [c3, s3], column rotations likewise.  Vector loads
        // complete in order, so several rays can be in flight and waited for
        // one at a time.  A ray at or past r1 gets a zero-sized buffer: its
        // loads return zeros, so its flux -- and its contribution -- is zero.
        struct RayRegs { v4u a; v2u b; v4u c; v2u d; v4u e; v2u f; };
        auto load = [&](RayRegs& d, int rr) {
            const bool live = rr < r1;
            const int q = live ? rr : r0;
            const __amdgpu_buffer_rsrc_t rp = __builtin_amdgcn_make_buffer_rsrc(
                (void*)(params + (int64_t)q * kRayStride), (short)0,
                live ? kRayStride * 4 : 0, kBufferFlags);
            const __amdgpu_buffer_rsrc_t rq = __builtin_amdgcn_make_buffer_rsrc(
                (void*)(rot + (int64_t)q * kRotStride), (short)0,
                live ? kRotStride * 4 : 0, kBufferFlags);
            d.a = __builtin_amdgcn_raw_buffer_load_b128(rp, 0, 0, 0);
            d.b = __builtin_amdgcn_raw_buffer_load_b64(rp, 16, 0, 0);
            d.c = __builtin_amdgcn_raw_buffer_load_b128(rq, 0, 0, 0);
            d.d = __builtin_amdgcn_raw_buffer_load_b64(rq, 16, 0, 0);
            d.e = __builtin_amdgcn_raw_buffer_load_b128(rq, 32, 0, 0);
            d.f = __builtin_amdgcn_raw_buffer_load_b64(rq, 48, 0, 0);
        };
        // operands of one ray from its parameters (block 0 evaluated, blocks
        // 1-3 rotated from it)
        auto operands = [&](const RayRegs& d, float (&av)[4], float (&bv)[CB]) {
            auto F = [](uint32_t u) { return __uint_as_float(u); };
            const float f = F(d.b.y);
            const float tha = step_phase(ci[0], F(d.a.x), F(d.a.y), F(d.b.x) + oa);
            const float as = f * sin_rev(tha), ac = f * cos_rev(tha);
            const float thb = step_phase(cj[0], F(d.a.z), F(d.a.w), ob);
            const float bs = sin_rev(thb), bc = cos_rev(thb);
            const float rc[3] = {F(d.c.x), F(d.c.z), F(d.d.x)};
            const float rs[3] = {F(d.c.y), F(d.c.w), F(d.d.y)};
            const float qc[3] = {F(d.e.x), F(d.e.z), F(d.f.x)};
            const float qs[3] = {F(d.e.y), F(d.e.w), F(d.f.y)};
            av[0] = as;
            bv[0] = bs;
            for (int k = 1; k < 4; k++) {
                av[k] = __builtin_fmaf(as, rc[k - 1], ac * rs[k - 1]);
                bv[k] = __builtin_fmaf(bs, qc[k - 1], bc * qs[k - 1]);
            }
        };
        // Software pipeline, one wave per SIMD.  Step r: the MFMAs of ray r
        // run while the VALU computes the operands of ray r + 1, interleaved
        // with them so that the VALU latencies hide under the MFMAs, and the
        // loads of ray r + 3 are issued (two steps of MFMA time to arrive).
        // Parameters rotate through three buffers and operands through two,
        // so the loop is unrolled by six and needs no register copies; the
        // scheduling barriers keep each stage in its step.  The step count
        // is rounded up to a multiple of six; the extra steps see rays past
        // r1, whose operands are zero.
        RayRegs pb[3];
        float av[2][4], bv[2][CB];
        auto step = [&](int r, int ia, const RayRegs& use, RayRegs& issue) {
            load(issue, r + 3);
            __builtin_amdgcn_sched_barrier(0);
            operands(use, av[ia ^ 1], bv[ia ^ 1]);
            for (int rb = 0; rb < 4; rb++)
                for (int cb = 0; cb < CB; cb++)
                    acc[rb][cb] = __builtin_amdgcn_mfma_f32_32x32x2f32(av[ia][rb], bv[ia][cb],
                                                                      acc[rb][cb], 0, 0, 0);
            for (int k = 0; k < 2 * CB; k++) {
                __builtin_amdgcn_sched_group_barrier(0x008, 2, 0);   // two MFMAs
                __builtin_amdgcn_sched_group_barrier(0x002, 4, 0);   // four VALU
            }
            __builtin_amdgcn_sched_barrier(0);
        };
        if (r0 < r1) {
            load(pb[0], r0);
            load(pb[1], r0 + 1);
            load(pb[2], r0 + 2);
            operands(pb[0], av[0], bv[0]);
            for (int r = r0; r < r1; r += 6) {
                step(r, 0, pb[1], pb[0]);
                step(r + 1, 1, pb[2], pb[1]);
                step(r + 2, 0, pb[0], pb[2]);
                step(r + 3, 1, pb[1], pb[0]);
                step(r + 4, 0, pb[2], pb[1]);
                step(r + 5, 1, pb[0], pb[2]);
            }
        }
    } else {
        // 32x32x16 bf16: lane supplies A'[l32][8*half + e], e = 0..7, i.e. rays
        // 4*half + e/2, part e%2; same for B'.
        // Each (part 0, part 1) pair of a ray is one dword of packed bf16 in
        // the operand, split as x = hi + lo by split_pair.  The phase
        // parameters of a group's rays for this lane (r + 4 half + q) are
        // buffer loads issued one group ahead of their use (two register sets
        // used in turn); the block rotations are loaded at the top of the
        // step and used after the phases are evaluated.  Rays at or past r1
        // read zeros, so their flux is zero.
        struct GroupRegs { v4u p[4]; v4u f[4]; };
        auto rsrc = [&](const float* base, int r, int stride) {
            // wave-uniform; read back as such (a clamp would otherwise be
            // computed on the VALU and the resource taken as divergent)
            const int nr = __builtin_amdgcn_readfirstlane(min(max(r1 - r, 0), kRayGroup));
            return __builtin_amdgcn_make_buffer_rsrc(
                (void*)(base + (int64_t)min(r, r1) * stride), (short)0, nr * stride * 4,
                kBufferFlags);
        };
        auto gload = [&](GroupRegs& g, int r) {
            const __amdgpu_buffer_rsrc_t rp = rsrc(params, r, kRayStride);
            for (int q = 0; q < 4; q++) {
                const int k = 4*half + q;
                g.p[q] = __builtin_amdgcn_raw_buffer_load_b128(rp, k * kRayStride * 4, 0, 0);
                g.f[q] = __builtin_amdgcn_raw_buffer_load_b128(rp, k * kRayStride * 4 + 16, 0, 0);
            }
        };
        auto op = [](const uint32_t (&w)[4]) {
            return __builtin_bit_cast(v8bf, (v4u){w[0], w[1], w[2], w[3]});
        };
        auto F = [](uint32_t u) { return __uint_as_float(u); };
        auto gstep = [&](int r, const GroupRegs& g, GroupRegs& next) {
            gload(next, r + kRayGroup);
            const __amdgpu_buffer_rsrc_t rq = rsrc(rot, r, kRotStride);
            v4u rc[4];
            v2u rd[4];
            for (int q = 0; q < 4; q++) {
                const int k = 4*half + q;
                rc[q] = __builtin_amdgcn_raw_buffer_load_b128(rq, k * kRotStride * 4, 0, 0);
                rd[q] = __builtin_amdgcn_raw_buffer_load_b64(rq, k * kRotStride * 4 + 16, 0, 0);
            }
            // column block 0 evaluated, blocks 1 .. CB-1 (32 k pixels on)
            // rotated from it
            v4u cc[4];
            v2u cd[4];
            if constexpr (CB > 2) {
                for (int q = 0; q < 4; q++) {
                    const int k = 4*half + q;
                    cc[q] = __builtin_amdgcn_raw_buffer_load_b128(rq, k * kRotStride * 4 + 32,
                                                                  0, 0);
                    cd[q] = __builtin_amdgcn_raw_buffer_load_b64(rq, k * kRotStride * 4 + 48,
                                                                 0, 0);
                }
            }
            uint32_t bhi[CB][4], blo[CB][4];                  // [block][ray q]
            for (int q = 0; q < 4; q++) {
                const float thb = step_phase(cj[0], F(g.p[q].z), F(g.p[q].w), 0.f);
                const float c = cos_rev(thb), sn = sin_rev(thb);
                split_pair(c, sn, bhi[0][q], blo[0][q]);
                for (int kb = 1; kb < CB; kb++) {
                    float cq, sq;
                    if constexpr (CB > 2) {
                        cq = F(kb == 3 ? cd[q].x : cc[q][2*kb - 2]);
                        sq = F(kb == 3 ? cd[q].y : cc[q][2*kb - 1]);
                    } else {
                        cq = F(g.f[q].z);
                        sq = F(g.f[q].w);
                    }
                    split_pair(__builtin_fmaf(c, cq, -(sn * sq)), __builtin_fmaf(sn, cq, c * sq),
                               bhi[kb][q], blo[kb][q]);
                }
            }
            // row block 0 evaluated, blocks 1-3 rotated from it; the lane's k
            // parts (v0, v1) of U = c + i s are (c, -s) on a real row and
            // (s, c) on an imaginary one, and both rotate as
            // (x, y) -> (x cr + y sr, y cr - x sr)
            float x[4], y[4];
            for (int q = 0; q < 4; q++) {
                const float th = step_phase(ci[0], F(g.p[q].x), F(g.p[q].y), F(g.f[q].x));
                const float f = F(g.f[q].y);
                const float fs0 = f * sin_rev(th), fc0 = f * cos_rev(th);
                x[q] = prow ? fs0 : fc0;
                y[q] = prow ? fc0 : -fs0;
            }
            for (int rb = 0; rb < 4; rb++) {
                uint32_t ahi[4], alo[4];
                for (int q = 0; q < 4; q++) {
                    if (rb == 0) {
                        split_pair(x[q], y[q], ahi[q], alo[q]);
                    } else {
                        const float cr = F(rb == 3 ? rd[q].x : rc[q][2*rb - 2]);
                        const float sr = F(rb == 3 ? rd[q].y : rc[q][2*rb - 1]);
                        split_pair(__builtin_fmaf(x[q], cr, y[q] * sr),
                                   __builtin_fmaf(y[q], cr, -(x[q] * sr)), ahi[q], alo[q]);
                    }
                }
                for (int cb = 0; cb < CB; cb++) {
                    v16f a = acc[rb][cb];
                    a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(op(alo), op(bhi[cb]), a, 0, 0, 0);
                    a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(op(ahi), op(blo[cb]), a, 0, 0, 0);
                    a = __builtin_amdgcn_mfma_f32_32x32x16_bf16(op(ahi), op(bhi[cb]), a, 0, 0, 0);
                    acc[rb][cb] = a;
                }
            }
        };
        // unrolled by two so the register sets alternate without copies; a
        // step past r1 reads zeros and adds nothing
        if (r0 < r1) {
            GroupRegs g0, g1;
            gload(g0, r0);
            for (int r = r0; r < r1; r += 2 * kRayGroup) {
                gstep(r, g0, g1);
                gstep(r + kRayGroup, g1, g0);
            }
        }
    }

    // D[row][col]: row = 8*(e/4) + 4*half + e%4, col = l32 (32x32 MFMA layout)
    // out[i][j][2] over the padded nxp x nyp image
#pragma unroll
    for (int rb = 0; rb < 4; rb++)
#pragma unroll
        for (int cb = 0; cb < CB; cb++)
#pragma unroll
            for (int e = 0; e < 16; e++) {
                const int row = 8*(e >> 2) + 4*half + (e & 3);
                const int i = i0 + rb*16 + (row >> 1);
                const int j = j0 + cb*32 + l32;
                out[((int64_t)i * nyp + j) * 2 + (row & 1)] = acc[rb][cb][e];
            }
}


template <int kPrec>
__global__ __launch_bounds__(gemm_threads<kPrec>(), gemm_waves_per_simd<kPrec>()) void k_psf_gemm(
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
__global__ __launch_bounds__(gemm_threads<kPrec>(), gemm_waves_per_simd<kPrec>()) void k_psf_gemm_batch(
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

namespace bt {

int gemm_args_ok(const char* name, const bt_rays& rays, const double* p0, const double* a1,
                 const double* a2, int nx, int ny, int precision, const double* amp_out)
{
    if (!rays_ok(rays) || !p0 || !a1 || !a2 || !amp_out)
        return fail(BT_E_NULL, std::string(name) + ": null argument");
    if (nx <= 0 || ny <= 0 || nx > 8192 || ny > 8192)
        return fail(BT_E_SIZE, std::string(name) + ": image size must be 1..8192 per side");
    if (rays.n > (int64_t)1 << 30)
        return fail(BT_E_SIZE, std::string(name) + ": more than 2^30 rays");
    if (precision != BT_PSF_FP32 && precision != BT_PSF_BF16X3)
        return fail(BT_E_KIND, std::string(name) + ": unknown precision");
    return 0;
}

// The launches of bt_huygens_gemm on a workspace of workspace_bytes bytes
// (arguments already checked).
int huygens_gemm_launch(bt_rays rays, const double* p0, const double* a1, const double* a2,
                        int nx, int ny, double t, int ignore_vignetted, int precision,
                        double* amp_out, void* workspace, int64_t workspace_bytes,
                        void* stream);

// ---- allocation-free variant --------------------------------------------
// bt_huygens_gemm_noalloc takes its scratch (the compacted per-ray
// parameters, per-block counts and offsets, and the K-split partial images
// of gemm_layout) from this fixed device array instead of a caller
// workspace.  Each GPU the library runs on holds its own copy until the
// process ends; calls on one GPU share it.
constexpr int64_t kNoAllocScratchBytes = (int64_t)128 << 20;
constexpr int kMaxDevices = 64;

std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

}  // namespace bt

static __device__ __align__(256) unsigned char scratch[bt::kNoAllocScratchBytes];

namespace bt {

// Device address of the current device's `scratch`, looked up on the first
// call on each device (with that device current) and kept: the host-side
// name `scratch` is a stand-in, never a device address.
int noalloc_scratch(void** out)
{
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return fail((int)e, "bt_huygens_gemm_noalloc: hipGetDevice failed");
    if (dev < 0 || dev >= kMaxDevices)
        return fail(BT_E_SIZE, "bt_huygens_gemm_noalloc: device index beyond 63");
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(e != hipSuccess ? (int)e : BT_E_NULL,
                        std::string("bt_huygens_gemm_noalloc: hipGetSymbolAddress(scratch) "
                                    "failed: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

}  // namespace bt

extern "C" int64_t bt_huygens_noalloc_scratch_bytes(void)
{
    return bt::kNoAllocScratchBytes;
}

extern "C" int bt_huygens_gemm_noalloc(bt_rays rays, const double* p0, const double* a1,
                                       const double* a2, int nx, int ny, double t,
                                       int ignore_vignetted, int precision, double* amp_out,
                                       void* stream)
{
    int rc = gemm_args_ok("bt_huygens_gemm_noalloc", rays, p0, a1, a2, nx, ny, precision,
                          amp_out);
    if (rc) return rc;
    const int64_t need = gemm_layout(rays.n, nx, ny).bytes;
    if (need > kNoAllocScratchBytes)
        return fail(BT_E_SIZE,
                    "bt_huygens_gemm_noalloc: " + std::to_string(rays.n) + " rays and a " +
                    std::to_string(nx) + " x " + std::to_string(ny) + " image need " +
                    std::to_string(need) + " bytes of scratch, beyond the fixed " +
                    std::to_string(kNoAllocScratchBytes) +
                    " (use bt_huygens_gemm with a workspace)");
    void* ws = nullptr;
    rc = noalloc_scratch(&ws);
    if (rc) return rc;
    return huygens_gemm_launch(rays, p0, a1, a2, nx, ny, t, ignore_vignetted, precision,
                               amp_out, ws, kNoAllocScratchBytes, stream);
}

extern "C" int bt_huygens_gemm(bt_rays rays, const double* p0, const double* a1,
                               const double* a2, int nx, int ny, double t,
                               int ignore_vignetted, int precision, double* amp_out,
                               void* workspace, int64_t workspace_bytes, void* stream)
{
    int rc = gemm_args_ok("bt_huygens_gemm", rays, p0, a1, a2, nx, ny, precision, amp_out);
    if (rc) return rc;
    if (!workspace) return fail(BT_E_NULL, "bt_huygens_gemm: null argument");
    return huygens_gemm_launch(rays, p0, a1, a2, nx, ny, t, ignore_vignetted, precision,
                               amp_out, workspace, workspace_bytes, stream);
}

int bt::huygens_gemm_launch(bt_rays rays, const double* p0, const double* a1,
                            const double* a2, int nx, int ny, double t, int ignore_vignetted,
                            int precision, double* amp_out, void* workspace,
                            int64_t workspace_bytes, void* stream)
{
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
        hipLaunchKernelGGL(k_psf_gemm<BT_PSF_FP32>, grid, dim3(gemm_threads<BT_PSF_FP32>()), 0, s, params, rot, total, nx, ny,
                           L.tiles_i, L.ksplit, L.nxp, L.nyp, partial);
    else
        hipLaunchKernelGGL(k_psf_gemm<BT_PSF_BF16X3>, grid, dim3(gemm_threads<BT_PSF_BF16X3>()), 0, s, params, rot, total, nx, ny,
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
        hipLaunchKernelGGL(k_psf_gemm_batch<BT_PSF_FP32>, dim3(grid),
                           dim3(gemm_threads<BT_PSF_FP32>()), 0, s, params,
                           rot, items, nitems, L.tiles_i, tiles, nx, ny, L.nxp, L.nyp, partial);
    else
        hipLaunchKernelGGL(k_psf_gemm_batch<BT_PSF_BF16X3>, dim3(grid),
                           dim3(gemm_threads<BT_PSF_BF16X3>()), 0, s, params,
                           rot, items, nitems, L.tiles_i, tiles, nx, ny, L.nxp, L.nyp, partial);
    rc = check_launch("bt_huygens_batch(mfma)");
    if (rc) return rc;
    const int64_t red_blocks = (((int64_t)nx * ny + kRedPix - 1) / kRedPix) * nseg;
    hipLaunchKernelGGL(k_psf_reduce_batch, dim3((unsigned)red_blocks), dim3(kBlock), 0, s, partial,
                       ranges, nseg, nx, ny, L.nxp, L.nyp, (int)L.cap, amp_out);
    return check_launch("bt_huygens_batch(reduce)");
}
