// bt_mlp.h -- the Newton-start network: 6 -> 64 -> 64 -> 1, ReLU.
//
// The network predicts, for a ray meeting a Newton surface, the residual
//   delta = dt* - dt_conic
// between the exact intersection time and the analytic root of the surface's
// base conic, from the ray's point on that conic and its direction (all in
// the surface's local frame).  It is trained in FP32 (bt_mlp.hip); inside the
// trace it is evaluated wave-cooperatively on MFMA: the two hidden layers as
// v_mfma_f32_32x32x16_f16 products with FP32 accumulation, laid out so each
// layer's accumulator registers are the next layer's B operand directly.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bt {

constexpr int kMlpIn = 6;
constexpr int kMlpHidden = 64;
// FP32 training parameter vector: W1[64][6], b1[64], W2[64][64], b2[64],
// w3[64], b3
constexpr int kMlpOffW1 = 0;
constexpr int kMlpOffB1 = kMlpOffW1 + kMlpHidden * kMlpIn;
constexpr int kMlpOffW2 = kMlpOffB1 + kMlpHidden;
constexpr int kMlpOffB2 = kMlpOffW2 + kMlpHidden * kMlpHidden;
constexpr int kMlpOffW3 = kMlpOffB2 + kMlpHidden;
constexpr int kMlpOffB3 = kMlpOffW3 + kMlpHidden;
constexpr int kMlpParams = kMlpOffB3 + 1;

typedef _Float16 h8 __attribute__((ext_vector_type(8)));
typedef float f16v __attribute__((ext_vector_type(16)));

// Device inference tables (built by bt_mlp_pack from the FP32 parameters):
//   norm: in_mean[6], in_inv_std[6], out_mean, out_std, b3, pad   (16 floats)
//   w1:   [2 rb][64 lanes][8] half  = [W1[n][0..5], b1[n], 0], n = 32 rb + lane%32
//   w2:   [2 rb2][4 s][64 lanes][8] half = W2[o][sigma(s, lane/32, e)],
//         o = 32 rb2 + lane%32, sigma(2 rb1 + t, h, e) = 32 rb1 + 16 t + 8 (e/4) + 4 h + e%4
//   b2:   [64] float; lane element e of row block rb2 reads
//         b2[32 rb2 + 8 (e/4) + 4 (lane/32) + e%4]
//   w3:   [64] float, read like b2
// 11,328 bytes: small enough that the trace keeps a plan's networks in LDS.
struct MlpPacked {
    float norm[16];
    _Float16 w1[2 * 64 * 8];
    _Float16 w2[2 * 4 * 64 * 8];
    float b2[kMlpHidden];
    float w3[kMlpHidden];
};

// Row of element e of a 32x32 MFMA accumulator held by lane half h.
__device__ __forceinline__ int acc_row(int e, int h) { return 8 * (e >> 2) + 4 * h + (e & 3); }

// Evaluate the network for the 64 rays of the calling wave (one per lane).
// Must be reached by all 64 lanes together (uniform control flow): MFMA
// operates on the whole wave.  `f` are the lane's raw features; returns the
// de-normalised prediction for the lane's own ray.
__device__ __forceinline__ float mlp_predict_wave(const MlpPacked* __restrict__ P,
                                                  const float (&f)[kMlpIn])
{
    const int lane = threadIdx.x & 63;
    const int h = lane >> 5;
    // normalised, augmented input (bias slot = 1) as 8 halves
    h8 x;
    for (int k = 0; k < kMlpIn; k++)
        x[k] = (_Float16)((f[k] - P->norm[k]) * P->norm[6 + k]);
    x[6] = (_Float16)1.0f;
    x[7] = (_Float16)0.0f;
    h8 zero;
    for (int k = 0; k < 8; k++) zero[k] = (_Float16)0.0f;
    float part[2] = {0.f, 0.f};
    const h8 b1c[2] = {(h == 0) ? x : zero, (h == 1) ? x : zero};
    // layer 1 for both column blocks (the rays of lane half 0 and of lane
    // half 1); relu(H1) in the accumulator layout is the layer-2 B operand.
    // Each weight fragment is loaded once and used for both blocks.
    h8 hb[2][2][2];   // [cb][rb1][t]
    for (int rb = 0; rb < 2; rb++) {
        const h8 a = *reinterpret_cast<const h8*>(P->w1 + (rb * 64 + lane) * 8);
        for (int cb = 0; cb < 2; cb++) {
            f16v acc;
            for (int e = 0; e < 16; e++) acc[e] = 0.f;
            acc = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b1c[cb], acc, 0, 0, 0);
            for (int t = 0; t < 2; t++)
                for (int e = 0; e < 8; e++) {
                    const float v = acc[8 * t + e];
                    hb[cb][rb][t][e] = (_Float16)(v > 0.f ? v : 0.f);
                }
        }
    }
    // layers 2 and 3, one output row block at a time
    for (int rb2 = 0; rb2 < 2; rb2++) {
        f16v acc0, acc1;
        const float* b2 = P->b2 + 32 * rb2;
        for (int e = 0; e < 16; e++) {
            const float b = b2[acc_row(e, h)];
            acc0[e] = b;
            acc1[e] = b;
        }
        for (int s = 0; s < 4; s++) {
            const h8 a = *reinterpret_cast<const h8*>(P->w2 + ((rb2 * 4 + s) * 64 + lane) * 8);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, hb[0][s >> 1][s & 1], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(a, hb[1][s >> 1][s & 1], acc1, 0, 0, 0);
        }
        const float* w3 = P->w3 + 32 * rb2;
        for (int e = 0; e < 16; e++) {
            const float w = w3[acc_row(e, h)];
            const float v0 = acc0[e], v1 = acc1[e];
            part[0] = __builtin_fmaf(w, v0 > 0.f ? v0 : 0.f, part[0]);
            part[1] = __builtin_fmaf(w, v1 > 0.f ? v1 : 0.f, part[1]);
        }
    }
    // the two lane halves hold the two halves of each column's output rows
    const float y0 = part[0] + __shfl_xor(part[0], 32);
    const float y1 = part[1] + __shfl_xor(part[1], 32);
    const float y = ((h == 0) ? y0 : y1) + P->norm[14];
    return y * P->norm[13] + P->norm[12];
}

}  // namespace bt
