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
// Each workgroup takes chunks of 64 rays: monomials and wavefront values to
// LDS, then the basis values (a 64-ray x J product with C, read through
// wave-uniform scalar loads), then every thread adds its share of the
// Gram entries over the chunk.  Per-workgroup partial Gram matrices are
// summed in workgroup order by a second kernel, so results do not depend on
// scheduling.
#include <hip/hip_runtime.h>
#include <string>

#include "bt_common.h"

namespace bt {

constexpr int kZChunk = 64;                                // rays per LDS chunk
constexpr int kZMaxN = 11;                                 // radial order
constexpr int kZMaxM = (kZMaxN + 1) * (kZMaxN + 2) / 2;    // 78 monomials
constexpr int kZMaxJ = 78;                                 // basis functions
constexpr int kZMaxEnt = (kZMaxJ + 1) * (kZMaxJ + 2) / 2;  // 3160 Gram entries
constexpr int kZThreads = 256;
constexpr int kZPer = (kZMaxEnt + kZThreads - 1) / kZThreads;   // 13 per thread
constexpr int kZGrid = 512;   // two 81 KB-LDS workgroups per CU

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

__global__ __launch_bounds__(kZThreads) void k_zernike_gram(ZArgs a)
{
    __shared__ double mono[kZMaxM][kZChunk];
    __shared__ double basis[kZMaxJ + 1][kZChunk + 1];   // +1: rows on distinct banks
    const int tid = threadIdx.x;
    const int J1 = a.J + 1;
    // this thread's Gram entries e = tid + k * 256 as (row, col), row <= col
    int er[kZPer], ec[kZPer];
    double acc[kZPer];
#pragma unroll
    for (int k = 0; k < kZPer; k++) {
        const int e = tid + k * kZThreads;
        int r = 0, rem = e;
        while (r < J1 && rem >= J1 - r) { rem -= J1 - r; r++; }
        er[k] = r;
        ec[k] = r + rem;
        acc[k] = 0.0;
    }
    for (int64_t base = (int64_t)blockIdx.x * kZChunk; base < a.n;
         base += (int64_t)gridDim.x * kZChunk) {
        const int64_t i = base + tid;
        const bool ok = tid < kZChunk && i < a.n && !(a.mask && a.mask[i]);
        // a chunk whose rays are all masked adds nothing (outside the pupil)
        if (!__syncthreads_or(ok)) continue;
        if (tid < kZChunk) {
            double u = 0.0, v = 0.0, w = 0.0;
            if (ok) {
                u = a.x[i] * a.inv_r;
                v = a.y[i] * a.inv_r;
                w = (a.t0 - a.t[i]) / a.wavelength;      // analysis.py:375, 639
                if (a.flux) w = w * a.flux[i];
            }
            double up[kZMaxN + 1], vp[kZMaxN + 1];
            up[0] = ok ? 1.0 : 0.0;
            vp[0] = 1.0;
#pragma unroll
            for (int k = 1; k <= kZMaxN; k++) {
                up[k] = up[k - 1] * u;
                vp[k] = vp[k - 1] * v;
            }
            // fully unrolled so up[] and vp[] stay in registers
#pragma unroll
            for (int d = 0; d <= kZMaxN; d++) {
                if (d <= a.nmax) {
#pragma unroll
                    for (int q = 0; q <= d; q++)
                        mono[d * (d + 1) / 2 + q][tid] = up[d - q] * vp[q];
                }
            }
            basis[a.J][tid] = w;
        }
        __syncthreads();
        for (int e = tid; e < a.J * kZChunk; e += kZThreads) {
            const int j = e / kZChunk, p = e % kZChunk;
            const double* cj = a.coef + (int64_t)j * a.M;
            // four interleaved partial sums (fixed order) shorten the
            // dependent FMA chain
            double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
            int m = 0;
            for (; m + 4 <= a.M; m += 4) {
                s0 += cj[m] * mono[m][p];
                s1 += cj[m + 1] * mono[m + 1][p];
                s2 += cj[m + 2] * mono[m + 2][p];
                s3 += cj[m + 3] * mono[m + 3][p];
            }
            for (; m < a.M; m++) s0 += cj[m] * mono[m][p];
            basis[j][p] = (s0 + s1) + (s2 + s3);
        }
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kZPer; k++) {
            if (tid + k * kZThreads < a.nent) {
                const double* br = basis[er[k]];
                const double* bc = basis[ec[k]];
                double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
                for (int p = 0; p < kZChunk; p += 4) {
                    s0 += br[p] * bc[p];
                    s1 += br[p + 1] * bc[p + 1];
                    s2 += br[p + 2] * bc[p + 2];
                    s3 += br[p + 3] * bc[p + 3];
                }
                acc[k] += (s0 + s1) + (s2 + s3);
            }
        }
        __syncthreads();
    }
    double* out = a.partial + (int64_t)blockIdx.x * a.nent;
#pragma unroll
    for (int k = 0; k < kZPer; k++) {
        const int e = tid + k * kZThreads;
        if (e < a.nent) out[e] = acc[k];
    }
}

// out[e] = sum over workgroups b of partial[b][e], in a fixed order: a
// block covers kZRedEnt consecutive entries (coalesced reads) with
// kZRedSlices lanes per entry, lane g summing b = g, g + kZRedSlices, ...;
// the slice sums are then added in g order.
constexpr int kZRedEnt = 32;
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
    hipLaunchKernelGGL(k_zernike_gram, dim3(kZGrid), dim3(kZThreads), 0, s, a);
    int rc = check_launch("bt_zernike_gram");
    if (rc) return rc;
    hipLaunchKernelGGL(k_zernike_reduce, dim3((a.nent + kZRedEnt - 1) / kZRedEnt), dim3(256), 0,
                       s, (const double*)a.partial, kZGrid, a.nent, gram);
    return check_launch("bt_zernike_gram");
}

}  // extern "C"
