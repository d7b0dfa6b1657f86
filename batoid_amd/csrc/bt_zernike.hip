// bt_zernike.hip -- Zernike decomposition of a wavefront on the GPU.
//
// batoid.zernike and batoid.zernikeGQ (batoid/analysis.py:479-641) fit or
// project the optical path difference of traced rays onto Zernike
// polynomials evaluated at the rays' entrance-pupil positions.  The O(N)
// part runs here, one pass over the rays: per ray the wavefront value
// w = (t0 - t) / wavelength (times the flux, for quadrature weights) and
// the J Zernike values Z_j(x/R, y/R), from the exact monomial expansions the
// host supplies (C[j][m], monomial m = x^(d-q) y^q at m = d(d+1)/2 + q).
// The output is the Gram matrix of the augmented basis [Z_1 .. Z_J, w]:
//   G[r][c] = sum over unmasked rays of B_r B_c,   r <= c <= J,
// which holds the least-squares normal matrix (r, c < J), the projections
// sum_i Z_j w (c = J) and sum_i w^2.  The host solves the J x J system.
//
// Each workgroup stages the coefficient table in LDS; each of its waves
// then takes groups of 16 rays on its own, and both products run on the
// FP64 matrix cores (v_mfma_f64_16x16x4f64: lane l supplies A[l%16][l/16]
// and B[l/16][l%16] and holds D[4e + l/16][l%16], e = 0..3):
//  1. lane l computes the powers of ray l % 16 and, in registers, the
//     monomials it supplies as A: 4 ks + l / 16 for each K step ks;
//  2. the basis values Z[p][j] = sum_m mono[p][m] C[j][m] of the 16 rays
//     against every 16-column block of C (K = monomials), written with the
//     wavefront values to the wave's own LDS rows;
//  3. the Gram blocks G[16 rb..][16 cb..] (rb <= cb) += sum_p B[p][r] B[p][c]
//     over the same 16 rays (K = rays), accumulated in registers.
// A wave only reads the LDS rows it wrote, so the ray loop has no barrier.
// At the end the four waves' Gram blocks are added in wave order in LDS
// and the workgroup writes its partial Gram matrix (upper triangle,
// row-major); a second kernel sums the partials in workgroup order, so the
// result does not depend on scheduling.  Groups whose rays are all masked
// are skipped (they add exact zeros).
#include <hip/hip_runtime.h>
#include <string>

#include "bt_common.h"

namespace bt {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int kZChunk = 64;                                // rays per LDS chunk
constexpr int kZMaxN = 11;                                 // radial order
constexpr int kZMaxM = (kZMaxN + 1) * (kZMaxN + 2) / 2;    // 78 monomials
constexpr int kZMaxJ = 78;                                 // basis functions
constexpr int kZThreads = 256;
constexpr int kZGrid = 512;

struct ZArgs {
    const double* x;
    const double* y;
    const double* t;
    const double* flux;       // NULL: weight 1
    const uint8_t* mask;      // NULL: every ray; else rays with mask != 0 skipped
    int64_t n;
    double t0, wavelength, inv_r;
    const double* coef;       // [J][M]
    int J, nmax, M, nent;
    double* partial;          // [gridDim.x][nent]
};

// K steps (4 monomials each) of the basis product for NB = ceil((J + 1) / 16):
// the largest J with that NB has radial order 4, 7, 9, 10, 11, i.e. 15, 36,
// 55, 66, 78 monomials
constexpr int zernike_ksteps(int nb)
{
    return nb == 1 ? 4 : nb == 2 ? 9 : nb == 3 ? 14 : nb == 4 ? 17 : 20;
}

// Radial order d and index q of monomial m = d(d+1)/2 + q (x^(d-q) y^q).
constexpr int mono_d(int m)
{
    int d = 0;
    while ((d + 1) * (d + 2) / 2 <= m) d++;
    return d;
}
constexpr int mono_q(int m) { return m - mono_d(m) * (mono_d(m) + 1) / 2; }

// Monomial m of a ray from its powers (0 past radial order kZMaxN).
template <int m>
__device__ __forceinline__ double mono_value(const double (&up)[kZMaxN + 1],
                                             const double (&vp)[kZMaxN + 1])
{
    constexpr int d = mono_d(m), q = mono_q(m);
    if constexpr (d > kZMaxN) {
        return 0.0;
    } else {
        return up[d - q] * vp[q];
    }
}

// The A operand of K step KSI for lane group lk: monomial 4 KSI + lk of the
// lane's ray, 0 at and past M.
template <int KSI>
__device__ __forceinline__ double mono_operand(const double (&up)[kZMaxN + 1],
                                               const double (&vp)[kZMaxN + 1], int lk, int M)
{
    const double c0 = mono_value<4 * KSI + 0>(up, vp);
    const double c1 = mono_value<4 * KSI + 1>(up, vp);
    const double c2 = mono_value<4 * KSI + 2>(up, vp);
    const double c3 = mono_value<4 * KSI + 3>(up, vp);
    const double v = lk == 0 ? c0 : lk == 1 ? c1 : lk == 2 ? c2 : c3;
    return 4 * KSI + lk < M ? v : 0.0;
}

// The basis product of K steps KSI .. KS-1 (unrolled; steps at and past
// ksteps add nothing and are skipped).
template <int NB, int KS, int KSI>
__device__ __forceinline__ void basis_steps(v4d (&z)[NB], const double (&up)[kZMaxN + 1],
                                            const double (&vp)[kZMaxN + 1], int lk, int li,
                                            int M, int ksteps, const double (*ct)[NB * 16])
{
    if constexpr (KSI < KS) {
        if (KSI < ksteps) {
            const double am = mono_operand<KSI>(up, vp, lk, M);
            const int m = KSI * 4 + lk;
#pragma unroll
            for (int jb = 0; jb < NB; jb++)
                z[jb] = __builtin_amdgcn_mfma_f64_16x16x4f64(am, ct[m][jb * 16 + li], z[jb], 0,
                                                             0, 0);
            basis_steps<NB, KS, KSI + 1>(z, up, vp, lk, li, M, ksteps, ct);
        }
    }
}

// NB = ceil((J + 1) / 16) Gram blocks per side.  Each wave works alone on
// groups of 16 rays (group g: rays 16 g .. 16 g + 15, the groups dealt to
// the grid's waves in turn): lane l handles ray l % 16 and computes, in
// registers, exactly the monomials it supplies as the basis product's A
// operand (4 ks + l / 16); the basis values go through the wave's own LDS
// rows to the Gram product.  No barrier inside the ray loop.
template <int NB>
__global__ __launch_bounds__(kZThreads) void k_zernike_gram(ZArgs a)
{
    constexpr int NBLK = NB * (NB + 1) / 2;
    constexpr int KS = zernike_ksteps(NB);
    constexpr int MP = KS * 4, JP = NB * 16;         // padded monomial and basis counts
    constexpr int kWaves = kZThreads / 64;
    __shared__ double ct[MP][JP];                    // C transposed, [monomial][basis j]
    __shared__ double basis[kWaves][16][JP];         // per wave: [ray][basis j]; j = J is w
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int li = lane & 15, lk = lane >> 4;
    const int J = a.J, J1 = J + 1, M = a.M;
    const int ksteps = (M + 3) / 4;
    for (int e = tid; e < MP * JP; e += kZThreads) {
        const int m = e / JP, j = e % JP;
        ct[m][j] = (j < J && m < M) ? a.coef[(int64_t)j * M + m] : 0.0;
    }
    // basis columns past J stay zero for every group
    double (*bw)[JP] = basis[wave];
    for (int e = lane; e < 16 * (JP - J1); e += 64) bw[e / (JP - J1)][J1 + e % (JP - J1)] = 0.0;
    __syncthreads();
    v4d g[NBLK];
    for (int b = 0; b < NBLK; b++) g[b] = (v4d){0.0, 0.0, 0.0, 0.0};
    struct In { double x, y, t, f; int keep; };
    auto fetch = [&](int64_t b, In& r) {
        const int64_t i = b + li;
        r.keep = 0;
        r.x = r.y = r.t = r.f = 0.0;
        if (i < a.n) {
            r.x = a.x[i];
            r.y = a.y[i];
            r.t = a.t[i];
            r.f = a.flux ? a.flux[i] : 1.0;
            r.keep = a.mask ? (a.mask[i] == 0) : 1;
        }
    };
    const int64_t stride = (int64_t)gridDim.x * kWaves * 16;
    int64_t base = ((int64_t)blockIdx.x * kWaves + wave) * 16;
    In cur, nxt;
    fetch(base, cur);
    for (; base < a.n; base += stride) {
        fetch(base + stride, nxt);                   // one group ahead
        const bool ok = cur.keep != 0;
        // a group whose rays are all masked adds nothing (outside the pupil)
        if (!__any(ok)) {
            cur = nxt;
            continue;
        }
        double u = 0.0, v = 0.0, w = 0.0;
        if (ok) {
            u = cur.x * a.inv_r;
            v = cur.y * a.inv_r;
            w = (a.t0 - cur.t) / a.wavelength;       // analysis.py:375, 639
            if (a.flux) w = w * cur.f;
        }
        double up[kZMaxN + 1], vp[kZMaxN + 1];
        up[0] = ok ? 1.0 : 0.0;
        vp[0] = 1.0;
#pragma unroll
        for (int k = 1; k <= kZMaxN; k++) {
            up[k] = up[k - 1] * u;
            vp[k] = vp[k - 1] * v;
        }
        // the previous group's Gram reads of these rows come first
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        if (lk == 0) bw[li][J] = w;
        // basis values of the group's 16 rays, every 16-column block of C
        v4d z[NB];
#pragma unroll
        for (int jb = 0; jb < NB; jb++) z[jb] = (v4d){0.0, 0.0, 0.0, 0.0};
        basis_steps<NB, KS, 0>(z, up, vp, lk, li, M, ksteps, ct);
#pragma unroll
        for (int jb = 0; jb < NB; jb++) {
            const int j = jb * 16 + li;
            if (j < J) {
#pragma unroll
                for (int e = 0; e < 4; e++) bw[4 * e + lk][j] = z[jb][e];
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        // Gram blocks over the group, 4 rays per step; the lane's value for
        // block row/column x is basis[ray][16 x + li]
#pragma unroll
        for (int ks = 0; ks < 4; ks++) {
            const int p = ks * 4 + lk;
            double xv[NB];
#pragma unroll
            for (int x = 0; x < NB; x++) xv[x] = bw[p][x * 16 + li];
            int b = 0;
#pragma unroll
            for (int rb = 0; rb < NB; rb++)
#pragma unroll
                for (int cb = rb; cb < NB; cb++, b++)
                    g[b] = __builtin_amdgcn_mfma_f64_16x16x4f64(xv[rb], xv[cb], g[b], 0, 0, 0);
        }
        cur = nxt;
    }
    // the four waves' blocks added in wave order, then the upper triangle out
    static_assert(NBLK * 256 <= kWaves * 16 * JP, "the wave sums fit in the basis buffer");
    double* sum = &basis[0][0][0];                   // NBLK x 256 doubles
    __syncthreads();
    for (int wv = 0; wv < kWaves; wv++) {
        if (wave == wv) {
#pragma unroll
            for (int b = 0; b < NBLK; b++)
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    double& d = sum[b * 256 + (4 * e + lk) * 16 + li];
                    d = (wv == 0) ? g[b][e] : d + g[b][e];
                }
        }
        __syncthreads();
    }
    double* out = a.partial + (int64_t)blockIdx.x * a.nent;
    for (int e = tid; e < a.nent; e += kZThreads) {
        int r = 0, rem = e;
        while (rem >= J1 - r) { rem -= J1 - r; r++; }
        const int c = r + rem;
        const int rb = r >> 4, cb = c >> 4;
        const int b = rb * NB - rb * (rb - 1) / 2 + (cb - rb);   // block (rb, cb), rb <= cb
        out[e] = sum[b * 256 + (r & 15) * 16 + (c & 15)];
    }
}

// out[e] = sum over workgroups b of partial[b][e], in a fixed order: a
// block covers kZRedEnt consecutive entries (coalesced reads) with
// kZRedSlices lanes per entry, lane g summing b = g, g + kZRedSlices, ...;
// the slice sums are then added in g order.
constexpr int kZRedEnt = 8;
constexpr int kZRedSlices = 256 / kZRedEnt;

__global__ __launch_bounds__(256) void k_zernike_reduce(const double* __restrict__ partial,
                                                        int nblocks, int nent,
                                                        double* __restrict__ out)
{
    __shared__ double part[kZRedSlices][kZRedEnt];
    const int pe = threadIdx.x % kZRedEnt, g = threadIdx.x / kZRedEnt;
    const int e = blockIdx.x * kZRedEnt + pe;
    double s = 0.0;
    if (e < nent)
        for (int b = g; b < nblocks; b += kZRedSlices) s += partial[(int64_t)b * nent + e];
    part[g][pe] = s;
    __syncthreads();
    if (g == 0 && e < nent) {
        double t = 0.0;
        for (int k = 0; k < kZRedSlices; k++) t += part[k][pe];
        out[e] = t;
    }
}

}  // namespace bt

using namespace bt;

extern "C" {

int64_t bt_zernike_workspace_bytes(int jmax)
{
    if (jmax < 1 || jmax > kZMaxJ) return -1;
    const int J1 = jmax + 1;
    return (int64_t)kZGrid * (J1 * (J1 + 1) / 2) * (int64_t)sizeof(double);
}

int bt_zernike_gram(const double* x, const double* y, const double* t, const double* flux,
                    const uint8_t* mask, int64_t n, double t0, double wavelength,
                    double r_outer, const double* coef, int jmax, int nmax, double* gram,
                    void* workspace, int64_t workspace_bytes, void* stream)
{
    if (jmax < 1 || jmax > kZMaxJ)
        return fail(BT_E_SIZE, "bt_zernike_gram: jmax must be in [1, 78]");
    if (nmax < 0 || nmax > kZMaxN)
        return fail(BT_E_SIZE, "bt_zernike_gram: nmax must be in [0, 11]");
    if (n < 0) return fail(BT_E_SIZE, "bt_zernike_gram: negative n");
    if (!coef || !gram || !workspace || (n > 0 && (!x || !y || !t)))
        return fail(BT_E_NULL, "bt_zernike_gram: null argument");
    if (!(r_outer > 0.0)) return fail(BT_E_SIZE, "bt_zernike_gram: r_outer must be > 0");
    const int64_t need = bt_zernike_workspace_bytes(jmax);
    if (workspace_bytes < need)
        return fail(BT_E_SIZE, "bt_zernike_gram: workspace too small (" +
                    std::to_string(need) + " bytes needed)");
    ZArgs a;
    a.x = x; a.y = y; a.t = t; a.flux = flux; a.mask = mask; a.n = n;
    a.t0 = t0; a.wavelength = wavelength; a.inv_r = 1.0 / r_outer;
    a.coef = coef; a.J = jmax; a.nmax = nmax;
    a.M = (nmax + 1) * (nmax + 2) / 2;
    a.nent = (jmax + 1) * (jmax + 2) / 2;
    a.partial = static_cast<double*>(workspace);
    hipStream_t s = as_stream(stream);
    // Gram blocks per side for jmax, raised if the table has more monomials
    // than that variant's K steps hold (a table of higher order than jmax needs)
    int nb = (jmax + 1 + 15) / 16;
    while (nb < 5 && zernike_ksteps(nb) < (a.M + 3) / 4) nb++;
    switch (nb) {
    case 1: hipLaunchKernelGGL(k_zernike_gram<1>, dim3(kZGrid), dim3(kZThreads), 0, s, a); break;
    case 2: hipLaunchKernelGGL(k_zernike_gram<2>, dim3(kZGrid), dim3(kZThreads), 0, s, a); break;
    case 3: hipLaunchKernelGGL(k_zernike_gram<3>, dim3(kZGrid), dim3(kZThreads), 0, s, a); break;
    case 4: hipLaunchKernelGGL(k_zernike_gram<4>, dim3(kZGrid), dim3(kZThreads), 0, s, a); break;
    default: hipLaunchKernelGGL(k_zernike_gram<5>, dim3(kZGrid), dim3(kZThreads), 0, s, a); break;
    }
    int rc = check_launch("bt_zernike_gram");
    if (rc) return rc;
    hipLaunchKernelGGL(k_zernike_reduce, dim3((a.nent + kZRedEnt - 1) / kZRedEnt), dim3(256), 0,
                       s, (const double*)a.partial, kZGrid, a.nent, gram);
    return check_launch("bt_zernike_gram");
}

}  // extern "C"
