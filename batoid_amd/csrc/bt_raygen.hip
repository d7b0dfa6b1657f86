// bt_raygen.hip -- parallel-beam ray generation on the GPU.
//
// One pass per RayVector: lane i computes ray i's stop coordinates (lattice,
// hexapolar ring, given points, or a counter-based draw from the PCG64 stream
// numpy's Generator would use), evaluates the stop surface's sag, moves the
// point into the ray frame, runs finishParallel's back-propagation and
// writes all 11 columns.  The arithmetic of each stage is the reference's
// (batoid/rayVector.py:219-643, 816-858; src/batoid.cpp:8-22, 40-69), in
// the same operation order, so lattice, random and given-point rays are
// bit-identical to the reference's; hexapolar rays differ from numpy's only
// through the device's FP64 sin/cos.
//
// The pass is a pure HBM write stream: 74 bytes per ray (9 FP64 columns
// and 2 flag bytes; POINTS also reads 16), no re-reads.  Measured at the
// rate of the same kernel with the trivial lattice pattern (~5.2 TB/s of
// writes), i.e. the ring lookup and sin/cos hide under the stores.
#include <hip/hip_runtime.h>
#include <string>

#include "bt_device.h"
#include "bt_common.h"

namespace bt {

// ---------------------------------------------------------------------------
// PCG64 (XSL-RR 128/64; O'Neill 2014) with arbitrary-stride jump-ahead
// (F. Brown, "Random Number Generation with Arbitrary Strides", 1994).
// ---------------------------------------------------------------------------
struct U128 {
    uint64_t hi, lo;
};

// Full 64 x 64 -> 128 product from 32-bit pieces (usable in constant
// evaluation, where the device intrinsic is not).
__host__ __device__ constexpr U128 mul_64x64(uint64_t a, uint64_t b)
{
    const uint64_t a0 = a & 0xffffffffull, a1 = a >> 32;
    const uint64_t b0 = b & 0xffffffffull, b1 = b >> 32;
    const uint64_t p00 = a0 * b0, p01 = a0 * b1, p10 = a1 * b0, p11 = a1 * b1;
    const uint64_t mid = (p00 >> 32) + (p01 & 0xffffffffull) + (p10 & 0xffffffffull);
    return U128{p11 + (p01 >> 32) + (p10 >> 32) + (mid >> 32),
                (mid << 32) | (p00 & 0xffffffffull)};
}

__host__ __device__ constexpr U128 mul_const(U128 a, U128 b)   // mod 2^128
{
    U128 p = mul_64x64(a.lo, b.lo);
    p.hi += a.lo * b.hi + a.hi * b.lo;
    return p;
}

__host__ __device__ constexpr U128 add128(U128 a, U128 b)
{
    const uint64_t lo = a.lo + b.lo;
    return U128{a.hi + b.hi + (lo < a.lo ? 1ull : 0ull), lo};
}

__device__ __forceinline__ U128 mul128(U128 a, U128 b)        // mod 2^128
{
    return U128{__umul64hi(a.lo, b.lo) + a.lo * b.hi + a.hi * b.lo, a.lo * b.lo};
}

// PCG's 128-bit LCG multiplier (pcg_variants.h, PCG_DEFAULT_MULTIPLIER_128).
constexpr U128 kPcgMult{2549297995355413924ull, 4865540595714422341ull};

// Jumps by 2^b steps: state -> A[b] state + inc G[b], with A[b] = M^(2^b)
// and G[b] = 1 + M + ... + M^(2^b - 1), both mod 2^128.
struct PcgJumps {
    U128 a[64];
    U128 g[64];
};

constexpr PcgJumps make_pcg_jumps()
{
    PcgJumps t{};
    U128 a = kPcgMult;
    U128 g{0, 1};
    for (int b = 0; b < 64; b++) {
        t.a[b] = a;
        t.g[b] = g;
        // doubling: M^(2k) = (M^k)^2, G_2k = G_k (M^k + 1)
        g = mul_const(g, add128(a, U128{0, 1}));
        a = mul_const(a, a);
    }
    return t;
}

__constant__ PcgJumps kPcgJumps = make_pcg_jumps();

// State after k steps from (state, inc): the k-th output's state.
__device__ __forceinline__ U128 pcg_state_at(U128 state, U128 inc, uint64_t k)
{
    U128 A{0, 1}, G{0, 0};
    for (int b = 0; k != 0; b++, k >>= 1) {
        if (k & 1) {
            // apply jump b after the accumulated one: A' = A_b A, G' = A_b G + G_b
            const U128 ab = kPcgJumps.a[b];
            G = add128(mul128(ab, G), kPcgJumps.g[b]);
            A = mul128(ab, A);
        }
    }
    return add128(mul128(A, state), mul128(inc, G));
}

__device__ __forceinline__ uint64_t pcg_output(U128 s)          // XSL-RR
{
    const uint64_t x = s.hi ^ s.lo;
    const unsigned rot = (unsigned)(s.hi >> 58);
    return (x >> rot) | (x << ((64u - rot) & 63u));
}

// The k-th (from 0) double of the stream: numpy's next_double,
// (next64 >> 11) * 2^-53, of the (k+1)-th step.
__device__ __forceinline__ double pcg_uniform(U128 state, U128 inc, uint64_t k)
{
    const uint64_t r = pcg_output(pcg_state_at(state, inc, k + 1));
    return (double)(r >> 11) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------------------
// The generation kernel
// ---------------------------------------------------------------------------
struct GenArgs {
    bt_raygen g;
    double vl[3];   // -vv in the finishParallel frame (batoid.cpp:44-46)
};

__device__ __forceinline__ int64_t gen_index()
{
    return (int64_t)xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
}

__global__ __launch_bounds__(kBlock) void k_generate(GenArgs a, bt_rays out)
{
    const int64_t i = gen_index();
    const int64_t n = out.n;
    if (i >= n) return;
    const bt_raygen& g = a.g;
    double xs, ys;
    switch (g.pattern) {
    case BT_GEN_POINTS:
        xs = g.px[i];
        ys = g.py[i];
        break;
    case BT_GEN_GRID:
    case BT_GEN_GRID_RANDOM: {
        double u, w;
        if (g.pattern == BT_GEN_GRID) {
            // fftshift(fftfreq(n, d))[k] = (k - n//2) / (n d)  (numpy.fft)
            const int64_t iy = i / g.nx, ix = i - iy * g.nx;
            u = (double)(ix - g.nx / 2) * g.fx;
            w = (double)(iy - g.ny / 2) * g.fy;
        } else {                                     // rayVector.py:415-416
            const U128 st{g.pcg_state[0], g.pcg_state[1]};
            const U128 inc{g.pcg_inc[0], g.pcg_inc[1]};
            u = g.lo[0] + g.span[0] * pcg_uniform(st, inc, (uint64_t)i);
            w = g.lo[1] + g.span[1] * pcg_uniform(st, inc, (uint64_t)(n + i));
        }
        xs = g.lx[0]*u + g.lx[1]*w;                  // rayVector.py:431-432
        ys = g.ly[0]*u + g.ly[1]*w;
        break;
    }
    case BT_GEN_POLAR:
    case BT_GEN_POLAR_RANDOM: {
        double rr, th;
        if (g.pattern == BT_GEN_POLAR) {
            // ring k: off[k] <= i < off[k+1]; lo stays in [0, nring)
            int lo = 0, hi = g.nring;
            while (hi - lo > 1) {
                const int mid = (lo + hi) >> 1;
                if (g.ring_off[mid] <= i) lo = mid; else hi = mid;
            }
            rr = g.ring_rho[lo];
            th = (double)(i - g.ring_off[lo]) * g.ring_dth[lo];   // linspace
        } else {                                     // rayVector.py:628-629
            const U128 st{g.pcg_state[0], g.pcg_state[1]};
            const U128 inc{g.pcg_inc[0], g.pcg_inc[1]};
            rr = sqrt(g.lo[0] + g.span[0] * pcg_uniform(st, inc, (uint64_t)i));
            th = g.lo[1] + g.span[1] * pcg_uniform(st, inc, (uint64_t)(n + i));
        }
        double sn, cs;
        sincos(th, &sn, &cs);
        xs = rr*cs;                                  // rayVector.py:630-631
        ys = rr*sn;
        break;
    }
    default:
        return;   // rejected on the host
    }
    const double zs = surf_sag(g.stop, xs, ys);
    // stop frame -> ray frame (batoid.cpp:8-22)
    const double* dr = g.stop_dr;
    const double* m = g.stop_drot;
    double dx = xs-dr[0];
    double dy = ys-dr[1];
    double dz = zs-dr[2];
    double x = dx*m[0] + dy*m[3] + dz*m[6];
    double y = dx*m[1] + dy*m[4] + dz*m[7];
    double z = dx*m[2] + dy*m[5] + dz*m[8];
    // finishParallel (batoid.cpp:40-69)
    const double* fr = g.origin;
    const double* f = g.rot;
    dx = x-fr[0];
    dy = y-fr[1];
    dz = z-fr[2];
    double xx = dx*f[0] + dy*f[3] + dz*f[6];
    double yy = dx*f[1] + dy*f[4] + dz*f[7];
    double zz = dx*f[2] + dy*f[5] + dz*f[8];
    const double dt = -zz/a.vl[2];
    xx += dt*a.vl[0];
    yy += dt*a.vl[1];
    zz += dt*a.vl[2];
    out.x[i] = xx*f[0] + yy*f[1] + zz*f[2] + fr[0];
    out.y[i] = xx*f[3] + yy*f[4] + zz*f[5] + fr[1];
    out.z[i] = xx*f[6] + yy*f[7] + zz*f[8] + fr[2];
    // constant columns (rayVector.py:831-841)
    out.vx[i] = g.vv[0];
    out.vy[i] = g.vv[1];
    out.vz[i] = g.vv[2];
    out.t[i] = 0.0;
    out.wavelength[i] = g.wavelength;
    out.flux[i] = g.flux ? g.flux[i] : g.flux0;
    out.vignetted[i] = 0;
    out.failed[i] = 0;
}

}  // namespace bt

using namespace bt;

extern "C" {

int bt_generate_rays(const bt_raygen* gen, bt_rays out, void* stream)
{
    if (!gen) return fail(BT_E_NULL, "bt_generate_rays: null generator");
    if (!rays_ok(out)) return fail(BT_E_NULL, "bt_generate_rays: incomplete ray columns");
    const bt_raygen& g = *gen;
    if (!g.stop) return fail(BT_E_NULL, "bt_generate_rays: null stop surface");
    switch (g.pattern) {
    case BT_GEN_POINTS:
        if (out.n > 0 && (!g.px || !g.py))
            return fail(BT_E_NULL, "bt_generate_rays: POINTS needs px, py");
        break;
    case BT_GEN_GRID:
        if (g.nx < 1 || g.ny < 1 || g.nx > out.n || g.nx * g.ny != out.n)
            return fail(BT_E_SIZE, "bt_generate_rays: GRID needs nx * ny == n");
        break;
    case BT_GEN_GRID_RANDOM:
    case BT_GEN_POLAR_RANDOM:
        break;
    case BT_GEN_POLAR:
        if (g.nring < 1) return fail(BT_E_SIZE, "bt_generate_rays: POLAR needs nring >= 1");
        if (!g.ring_off || !g.ring_rho || !g.ring_dth)
            return fail(BT_E_NULL, "bt_generate_rays: POLAR needs the ring tables");
        break;
    default:
        return fail(BT_E_KIND, "bt_generate_rays: unknown pattern");
    }
    if (out.n == 0) return 0;
    GenArgs a;
    a.g = g;
    // batoid.cpp:44-46, evaluated once on the host as the reference does
    const double* vv = g.vv;
    const double* d = g.rot;
    a.vl[0] = -vv[0]*d[0] - vv[1]*d[3] - vv[2]*d[6];
    a.vl[1] = -vv[0]*d[1] - vv[1]*d[4] - vv[2]*d[7];
    a.vl[2] = -vv[0]*d[2] - vv[1]*d[5] - vv[2]*d[8];
    hipLaunchKernelGGL(k_generate, dim3(grid_for(out.n)), dim3(kBlock), 0, as_stream(stream),
                       a, out);
    return check_launch("bt_generate_rays");
}

}  // extern "C"
