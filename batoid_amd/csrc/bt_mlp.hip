// bt_mlp.hip -- training and inference kernels of the Newton-start network,
// and the sampler that draws its training data from the tracer itself.
//
// Training is FP32 on the vector units: one sample per lane, forward and
// backward through 6 -> 64 -> 64 -> 1 (ReLU), per-block gradient sums in
// LDS, a fixed-order reduction over blocks (run-to-run deterministic), and
// an Adam step.  Inference inside the trace uses the packed fp16/MFMA form
// of bt_mlp.h.
#include <hip/hip_runtime.h>
#include <string>

#include "bt_device.h"
#include "bt_common.h"
#include "bt_mlp.h"

namespace bt {

constexpr int kMlpBlock = 64;   // samples (and lanes) per training block

__device__ __forceinline__ float relu(float v) { return v > 0.f ? v : 0.f; }

// Loss and per-block gradient sums for one minibatch.
//   loss = mean over the batch of (y - target)^2, y in normalised units
// Activations live in LDS rows padded to 65 floats, so a lane reading its
// own row and all lanes reading one column are both free of bank conflicts;
// weights are wave-uniform (scalar) loads.
constexpr int kRow = kMlpHidden + 1;

__global__ __launch_bounds__(kMlpBlock) void k_mlp_fwdbwd(
    const float* __restrict__ P, const float* __restrict__ X,
    const float* __restrict__ T, const int* __restrict__ idx, int64_t batch,
    const float* __restrict__ norm, float* __restrict__ part)
{
    __shared__ float s_h1[kMlpBlock][kRow];
    __shared__ float s_h2[kMlpBlock][kRow];
    __shared__ float s_g2[kMlpBlock][kRow];     // dL/d(pre-activation 2)
    __shared__ float s_g1[kMlpBlock][kRow];     // dL/d(pre-activation 1)
    __shared__ float s_x[kMlpBlock][kMlpIn];
    __shared__ float s_dy[kMlpBlock];
    __shared__ float s_e2[kMlpBlock];

    const int t = threadIdx.x;
    const int64_t b = (int64_t)blockIdx.x * kMlpBlock + t;
    const bool live = b < batch;
    const int64_t src = live ? (idx ? (int64_t)idx[b] : b) : 0;

    float x[kMlpIn];
    for (int k = 0; k < kMlpIn; k++)
        x[k] = live ? (X[src * kMlpIn + k] - norm[k]) * norm[kMlpIn + k] : 0.f;
    for (int j = 0; j < kMlpHidden; j++) {
        float a = P[kMlpOffB1 + j];
        for (int k = 0; k < kMlpIn; k++) a = __builtin_fmaf(P[kMlpOffW1 + j * kMlpIn + k], x[k], a);
        s_h1[t][j] = relu(a);
    }
    float y = P[kMlpOffB3];
    for (int k = 0; k < kMlpHidden; k++) {
        float a = P[kMlpOffB2 + k];
        for (int j = 0; j < kMlpHidden; j++)
            a = __builtin_fmaf(P[kMlpOffW2 + k * kMlpHidden + j], s_h1[t][j], a);
        const float h2 = relu(a);
        s_h2[t][k] = h2;
        y = __builtin_fmaf(P[kMlpOffW3 + k], h2, y);
    }
    const float err = live ? y - T[src] : 0.f;
    const float dy = 2.f * err / (float)batch;
    for (int k = 0; k < kMlpHidden; k++)
        s_g2[t][k] = s_h2[t][k] > 0.f ? dy * P[kMlpOffW3 + k] : 0.f;
    for (int j = 0; j < kMlpHidden; j++) {
        float a = 0.f;
        for (int k = 0; k < kMlpHidden; k++)
            a = __builtin_fmaf(s_g2[t][k], P[kMlpOffW2 + k * kMlpHidden + j], a);
        s_g1[t][j] = s_h1[t][j] > 0.f ? a : 0.f;
    }
    for (int k = 0; k < kMlpIn; k++) s_x[t][k] = x[k];
    s_dy[t] = dy;
    s_e2[t] = err * err;
    __syncthreads();

    // block sums: thread t owns column t of every weight matrix
    float* out = part + (int64_t)blockIdx.x * kMlpParams;
    for (int k = 0; k < kMlpHidden; k++) {             // dW2[k][t]
        float a = 0.f;
        for (int s = 0; s < kMlpBlock; s++) a = __builtin_fmaf(s_g2[s][k], s_h1[s][t], a);
        out[kMlpOffW2 + k * kMlpHidden + t] = a;
    }
    float gb2 = 0.f, gw3 = 0.f, gb1 = 0.f, gw1[kMlpIn] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int s = 0; s < kMlpBlock; s++) {
        gb2 += s_g2[s][t];
        gw3 = __builtin_fmaf(s_dy[s], s_h2[s][t], gw3);
        gb1 += s_g1[s][t];
        for (int k = 0; k < kMlpIn; k++) gw1[k] = __builtin_fmaf(s_g1[s][t], s_x[s][k], gw1[k]);
    }
    out[kMlpOffB2 + t] = gb2;
    out[kMlpOffW3 + t] = gw3;
    out[kMlpOffB1 + t] = gb1;
    for (int k = 0; k < kMlpIn; k++) out[kMlpOffW1 + t * kMlpIn + k] = gw1[k];
    if (t == 0) {
        float d = 0.f, e = 0.f;
        for (int s = 0; s < kMlpBlock; s++) { d += s_dy[s]; e += s_e2[s]; }
        out[kMlpOffB3] = d;
        part[(int64_t)gridDim.x * kMlpParams + blockIdx.x] = e;   // loss partials after the grads
    }
}

__global__ __launch_bounds__(kBlock) void k_mlp_reduce(const float* __restrict__ part, int nb,
                                                       float* __restrict__ grad,
                                                       float* __restrict__ loss, int64_t batch)
{
    const int p = blockIdx.x * kBlock + threadIdx.x;
    if (p < kMlpParams) {
        float a = 0.f;
        for (int b = 0; b < nb; b++) a += part[(int64_t)b * kMlpParams + p];
        grad[p] = a;
    } else if (p == kMlpParams && loss) {
        float e = 0.f;
        for (int b = 0; b < nb; b++) e += part[(int64_t)nb * kMlpParams + b];
        *loss = e / (float)batch;
    }
}

__global__ __launch_bounds__(kBlock) void k_mlp_adam(float* __restrict__ P,
                                                     const float* __restrict__ g,
                                                     float* __restrict__ m, float* __restrict__ v,
                                                     int64_t n, float lr, float b1, float b2,
                                                     float eps, float c1, float c2)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    P[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

// FP32 inference in normalised output units.
__global__ __launch_bounds__(kBlock) void k_mlp_forward(const float* __restrict__ P,
                                                        const float* __restrict__ norm,
                                                        const float* __restrict__ X, int64_t n,
                                                        float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    float x[kMlpIn];
    for (int k = 0; k < kMlpIn; k++) x[k] = (X[i * kMlpIn + k] - norm[k]) * norm[kMlpIn + k];
    float h1[kMlpHidden];
    for (int j = 0; j < kMlpHidden; j++) {
        float a = P[kMlpOffB1 + j];
        for (int k = 0; k < kMlpIn; k++) a = __builtin_fmaf(P[kMlpOffW1 + j * kMlpIn + k], x[k], a);
        h1[j] = relu(a);
    }
    float y = P[kMlpOffB3];
    for (int k = 0; k < kMlpHidden; k++) {
        float a = P[kMlpOffB2 + k];
        for (int j = 0; j < kMlpHidden; j++) a = __builtin_fmaf(P[kMlpOffW2 + k * kMlpHidden + j], h1[j], a);
        y = __builtin_fmaf(P[kMlpOffW3 + k], relu(a), y);
    }
    out[i] = y;
}

// FP32 parameters -> packed fp16 MFMA tables (one block of 64 threads).
__global__ __launch_bounds__(64) void k_mlp_pack(const float* __restrict__ P,
                                                 const float* __restrict__ norm,
                                                 MlpPacked* __restrict__ Q)
{
    const int lane = threadIdx.x;
    const int h = lane >> 5, l32 = lane & 31;
    if (lane < 16) Q->norm[lane] = lane < 12 ? norm[lane] : 0.f;
    if (lane == 12) Q->norm[12] = norm[12];
    if (lane == 13) Q->norm[13] = norm[13];
    if (lane == 14) Q->norm[14] = P[kMlpOffB3];
    for (int rb = 0; rb < 2; rb++) {
        const int n = 32 * rb + l32;
        _Float16* w1 = Q->w1 + (rb * 64 + lane) * 8;
        for (int k = 0; k < kMlpIn; k++) w1[k] = (_Float16)P[kMlpOffW1 + n * kMlpIn + k];
        w1[6] = (_Float16)P[kMlpOffB1 + n];
        w1[7] = (_Float16)0.f;
    }
    for (int rb2 = 0; rb2 < 2; rb2++) {
        const int o = 32 * rb2 + l32;
        for (int s = 0; s < 4; s++) {
            _Float16* w2 = Q->w2 + ((rb2 * 4 + s) * 64 + lane) * 8;
            const int rb1 = s >> 1, t = s & 1;
            for (int e = 0; e < 8; e++) {
                const int in = 32 * rb1 + 16 * t + 8 * (e >> 2) + 4 * h + (e & 3);
                w2[e] = (_Float16)P[kMlpOffW2 + o * kMlpHidden + in];
            }
        }
    }
    Q->b2[lane] = P[kMlpOffB2 + lane];
    Q->w3[lane] = P[kMlpOffW3 + lane];
}

// Packed (MFMA) inference, de-normalised; whole waves, padded lanes masked.
__global__ __launch_bounds__(256) void k_mlp_forward_packed(const MlpPacked* __restrict__ Q,
                                                            const float* __restrict__ X, int64_t n,
                                                            float* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const bool live = i < n;
    float f[kMlpIn];
    for (int k = 0; k < kMlpIn; k++) f[k] = live ? X[i * kMlpIn + k] : Q->norm[k];
    const float y = mlp_predict_wave(Q, f);
    if (live) out[i] = y;
}

struct SampleOut {
    float* feat;       // [n][6]: point on the base conic, direction
    double* dt_conic;  // [n]
    double* dt_exact;  // [n]
    uint8_t* status;   // bit 0: base start ok, bit 1: reference intersection ok
};

__global__ __launch_bounds__(kBlock) void k_newton_samples(Xform xf, const bt_surface* surf,
                                                           bt_rays rv, int niter, SampleOut o)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= rv.n) return;
    const double* dr = xf.dr;
    const double* m = xf.drot;
    double dx = rv.x[i]-dr[0], dy = rv.y[i]-dr[1], dz = rv.z[i]-dr[2];
    double x = dx*m[0] + dy*m[3] + dz*m[6];
    double y = dx*m[1] + dy*m[4] + dz*m[7];
    double z = dx*m[2] + dy*m[5] + dz*m[8];
    double vx = rv.vx[i]*m[0] + rv.vy[i]*m[3] + rv.vz[i]*m[6];
    double vy = rv.vx[i]*m[1] + rv.vy[i]*m[4] + rv.vz[i]*m[7];
    double vz = rv.vx[i]*m[2] + rv.vy[i]*m[5] + rv.vz[i]*m[8];
    uint8_t st = 0;
    double dtc = 0.0, dte = 0.0;
    if (!rv.failed[i]) {
        if (conic_start(surf, x, y, z, vx, vy, vz, dtc)) st |= 1;
        if (surf_tti(surf, x, y, z, vx, vy, vz, dte, niter)) st |= 2;
    }
    float* f = o.feat + i * kMlpIn;
    f[0] = (float)(x + vx*dtc);
    f[1] = (float)(y + vy*dtc);
    f[2] = (float)(z + vz*dtc);
    f[3] = (float)vx;
    f[4] = (float)vy;
    f[5] = (float)vz;
    o.dt_conic[i] = dtc;
    o.dt_exact[i] = dte;
    o.status[i] = st;
}

}  // namespace bt

using namespace bt;

extern "C" {

int bt_mlp_param_count(void) { return kMlpParams; }

int64_t bt_mlp_packed_bytes(void) { return (int64_t)sizeof(MlpPacked); }

int64_t bt_mlp_grad_workspace_bytes(int64_t batch)
{
    const int64_t nb = (batch + kMlpBlock - 1) / kMlpBlock;
    return (nb * kMlpParams + nb) * (int64_t)sizeof(float);
}

int bt_mlp_loss_grad(const float* params, const float* features, const float* targets,
                     const int32_t* index, int64_t batch, const float* in_norm, float* grad,
                     float* loss, void* workspace, int64_t workspace_bytes, void* stream)
{
    if (!params || !features || !targets || !in_norm || !grad || !workspace)
        return fail(BT_E_NULL, "bt_mlp_loss_grad: null argument");
    if (batch <= 0 || batch > ((int64_t)1 << 31)) return fail(BT_E_SIZE, "bt_mlp_loss_grad: bad batch");
    if (workspace_bytes < bt_mlp_grad_workspace_bytes(batch))
        return fail(BT_E_SIZE, "bt_mlp_loss_grad: workspace too small");
    const int nb = (int)((batch + kMlpBlock - 1) / kMlpBlock);
    float* part = static_cast<float*>(workspace);
    hipLaunchKernelGGL(k_mlp_fwdbwd, dim3(nb), dim3(kMlpBlock), 0, as_stream(stream), params,
                       features, targets, index, batch, in_norm, part);
    int rc = check_launch("bt_mlp_loss_grad");
    if (rc) return rc;
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + 1 + kBlock - 1) / kBlock), dim3(kBlock), 0,
                       as_stream(stream), part, nb, grad, loss, batch);
    return check_launch("bt_mlp_loss_grad(reduce)");
}

int bt_mlp_adam(float* params, const float* grad, float* m, float* v, int64_t nparam, float lr,
                float beta1, float beta2, float eps, int step, void* stream)
{
    if (!params || !grad || !m || !v) return fail(BT_E_NULL, "bt_mlp_adam: null argument");
    if (nparam <= 0 || step < 1) return fail(BT_E_SIZE, "bt_mlp_adam: bad size or step");
    const float c1 = 1.f - powf(beta1, (float)step);
    const float c2 = 1.f - powf(beta2, (float)step);
    hipLaunchKernelGGL(k_mlp_adam, dim3(grid_for(nparam)), dim3(kBlock), 0, as_stream(stream),
                       params, grad, m, v, nparam, lr, beta1, beta2, eps, c1, c2);
    return check_launch("bt_mlp_adam");
}

int bt_mlp_forward(const float* params, const float* in_norm, const float* features, int64_t n,
                   float* out, void* stream)
{
    if (!params || !in_norm || (n > 0 && (!features || !out)))
        return fail(BT_E_NULL, "bt_mlp_forward: null argument");
    if (n < 0) return fail(BT_E_SIZE, "bt_mlp_forward: negative n");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_mlp_forward, dim3(grid_for(n)), dim3(kBlock), 0, as_stream(stream),
                       params, in_norm, features, n, out);
    return check_launch("bt_mlp_forward");
}

int bt_mlp_pack(const float* params, const float* norm, void* packed, void* stream)
{
    if (!params || !norm || !packed) return fail(BT_E_NULL, "bt_mlp_pack: null argument");
    hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(64), 0, as_stream(stream), params, norm,
                       static_cast<MlpPacked*>(packed));
    return check_launch("bt_mlp_pack");
}

int bt_mlp_forward_packed(const void* packed, const float* features, int64_t n, float* out,
                          void* stream)
{
    if (!packed || (n > 0 && (!features || !out)))
        return fail(BT_E_NULL, "bt_mlp_forward_packed: null argument");
    if (n < 0) return fail(BT_E_SIZE, "bt_mlp_forward_packed: negative n");
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_mlp_forward_packed, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       as_stream(stream), static_cast<const MlpPacked*>(packed), features, n, out);
    return check_launch("bt_mlp_forward_packed");
}

int bt_newton_samples(const bt_surface* surface, const double* dr, const double* drot,
                      bt_rays rays, int niter, float* features, double* dt_conic,
                      double* dt_exact, uint8_t* status, void* stream)
{
    if (!surface || !dr || !drot || !rays_ok(rays) ||
        (rays.n > 0 && (!features || !dt_conic || !dt_exact || !status)))
        return fail(BT_E_NULL, "bt_newton_samples: null argument");
    if (niter < 0) return fail(BT_E_SIZE, "bt_newton_samples: negative niter");
    if (rays.n == 0) return 0;
    SampleOut o{features, dt_conic, dt_exact, status};
    hipLaunchKernelGGL(k_newton_samples, dim3(grid_for(rays.n)), dim3(kBlock), 0,
                       as_stream(stream), make_xform(dr, drot), surface, rays, niter, o);
    return check_launch("bt_newton_samples");
}

}  // extern "C"
