// closure.h — learned constitutive closure: a 4-64-64-3 SiLU network mapping the SVD invariants of the
// elastic deformation gradient to the Kirchhoff stress of jelly and snow.
//
// Inputs (normalised): x0 = (s1 - 1.1) / 0.5, x1 = (s2 - 1.1) / 0.5 with s1 >= s2 the singular values of F,
// x2 = theta / (pi/2) with theta in (-pi/2, pi/2] the angle of U (F = U diag(s) V^T), x3 = (Jh - 1) / 0.3
// with Jh the plastic volume ratio Jp for snow and 1 for jelly.  Outputs: tau / mu, (tau00, tau01, tau11).
// Trained range: s1, s2 in [0.6, 1.6], Jh in [0.7, 1.3]; outside it the analytic fixed-corotated /
// snow stress is used.
//
// Parameter layout (float, row = output unit): W1[64][4] | b1[64] | W2[64][64] | b2[64] | W3[3][64] | b3[3].
#pragma once

#include "mpm_device.h"

namespace mpm {

constexpr int kHid = 64;
constexpr int kOffW1 = 0;
constexpr int kOffB1 = 256;
constexpr int kOffW2 = 320;
constexpr int kOffB2 = 4416;
constexpr int kOffW3 = 4480;
constexpr int kOffB3 = 4672;
constexpr int kNumWeights = 4675;
constexpr float kSigLo = 0.6f, kSigHi = 1.6f, kJhLo = 0.7f, kJhHi = 1.3f;

typedef __attribute__((__vector_size__(16 * sizeof(float)))) float floatx16;

__device__ __forceinline__ float silu(float z)
{
    // z * sigmoid(z) with the hardware exp2 and reciprocal
    return z * __builtin_amdgcn_rcpf(1.f + __builtin_amdgcn_exp2f(-1.44269504f * z));
}

// Closure features of one particle; returns false when outside the trained range (then x is unspecified).
__device__ __forceinline__ bool closure_features(const Mat2& F, float jh, float x[4])
{
    float uc, us, s0, s1, vc, vs;
    svd2(F, uc, us, s0, s1, vc, vs);
    // canonical order s0 >= s1: swapping the singular values rotates U by 90 degrees (tau is unchanged)
    float th = atan2f(us, uc);
    if (s1 > s0) {
        const float t = s0;
        s0 = s1;
        s1 = t;
        th += 1.57079632679f;
    }
    // tau = U diag U^T is invariant under U -> -U: reduce the angle to (-pi/2, pi/2]
    if (th > 1.57079632679f) { th -= 3.14159265359f; }
    if (th <= -1.57079632679f) { th += 3.14159265359f; }
    x[0] = (s0 - 1.1f) * 2.f;
    x[1] = (s1 - 1.1f) * 2.f;
    x[2] = th * 0.636619772368f;
    x[3] = (jh - 1.f) * (1.f / 0.3f);
    return s0 >= kSigLo && s0 <= kSigHi && s1 >= kSigLo && s1 <= kSigHi && jh >= kJhLo && jh <= kJhHi;
}

// LDS staging of the small layers (per work group): W1 rows as float4, b1, b2, W3 columns as float4
// (W3[0][u], W3[1][u], W3[2][u], 0), b3.
struct ClosureLds {
    float4 w1[kHid];
    float b1[kHid];
    float b2[kHid];
    float4 w3[kHid];
    float b3[4];
    bf16x8 w2[8 * 64];
};

__device__ __forceinline__ void closure_stage(const ClosureDev& c, ClosureLds* s)
{
    for (int u = threadIdx.x; u < kHid; u += blockDim.x) {
        s->w1[u] = make_float4(c.w[kOffW1 + 4 * u], c.w[kOffW1 + 4 * u + 1], c.w[kOffW1 + 4 * u + 2],
                               c.w[kOffW1 + 4 * u + 3]);
        s->b1[u] = c.w[kOffB1 + u];
        s->b2[u] = c.w[kOffB2 + u];
        s->w3[u] = make_float4(c.w[kOffW3 + u], c.w[kOffW3 + kHid + u], c.w[kOffW3 + 2 * kHid + u], 0.f);
    }
    if (threadIdx.x < 4) { s->b3[threadIdx.x] = threadIdx.x < 3 ? c.w[kOffB3 + threadIdx.x] : 0.f; }
    for (int i = threadIdx.x; i < 8 * 64; i += blockDim.x) { s->w2[i] = c.w2_mfma[i]; }
}

__device__ __forceinline__ __bf16 to_bf16(float f) { return (__bf16)f; }

// Network output (tau / mu) of the particle of every lane of the wave.  Must be called by all 64 lanes
// (wave-uniform control flow); x holds the lane's own features.
//
// Layout (v_mfma_f32_32x32x16_bf16, checked on the device): A lane l holds A[l%32][8(l/32)+j],
// B lane l holds B[8(l/32)+j][l%32], C lane l holds C[8(r/4)+4(l/32)+r%4][l%32].  Particles are the N
// dimension: N-tile nt holds particles 32 nt + (l%32).  Layer 1 runs on the VALU directly in the B layout
// (each lane evaluates the 32 hidden units it feeds to the MFMA, for both N-tiles); layer 2 is 2 M-tiles x
// 4 K-steps of MFMA per N-tile with FP32 accumulation; layer 3 (FP32, VALU) reduces the 32 units a lane
// holds, and the two half-waves' partial sums are combined with permlane32 swaps.
__device__ __forceinline__ void closure_eval_wave(const ClosureLds* s, const float x[4], float y[3])
{
    const int lane = threadIdx.x & 63;
    const int half = lane >> 5;
    float part[2][3];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        // features of particle 32 nt + lane % 32 (s[0] = value of lane l%32, s[1] = value of lane l%32 + 32)
        float f[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(x[k]), __float_as_uint(x[k]), false,
                                                             false);
            f[k] = __uint_as_float(nt == 0 ? sw[0] : sw[1]);
        }
        // layer 1 for the units this lane feeds: u = 16 ks + 8 half + j
        bf16x8 bop[4];
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int u = 16 * ks + 8 * half + j;
                const float4 w = s->w1[u];
                float z = s->b1[u];
                z = fmaf(w.x, f[0], z);
                z = fmaf(w.y, f[1], z);
                z = fmaf(w.z, f[2], z);
                z = fmaf(w.w, f[3], z);
                bop[ks][j] = to_bf16(silu(z));
            }
        }
        // layer 2: z2[unit][particle] = sum_k W2[unit][k] h1[k][particle]
        float p0 = 0.f, p1 = 0.f, p2 = 0.f;
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            floatx16 acc = {};
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(s->w2[(mt * 4 + ks) * 64 + lane], bop[ks], acc, 0, 0, 0);
            }
            // layer 3 partial sums over the 16 units of this M-tile held by the lane
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int u = 32 * mt + 8 * (r >> 2) + 4 * half + (r & 3);
                const float h = silu(acc[r] + s->b2[u]);
                const float4 w3 = s->w3[u];
                p0 = fmaf(w3.x, h, p0);
                p1 = fmaf(w3.y, h, p1);
                p2 = fmaf(w3.z, h, p2);
            }
        }
        part[nt][0] = p0;
        part[nt][1] = p1;
        part[nt][2] = p2;
    }
    // the lane's own particle is in N-tile `half`; add the other half-wave's partial sum for it
#pragma unroll
    for (int o = 0; o < 3; ++o) {
        const auto s0 = __builtin_amdgcn_permlane32_swap(__float_as_uint(part[0][o]), __float_as_uint(part[0][o]),
                                                         false, false);
        const auto s1 = __builtin_amdgcn_permlane32_swap(__float_as_uint(part[1][o]), __float_as_uint(part[1][o]),
                                                         false, false);
        const float own = half == 0 ? part[0][o] : part[1][o];
        const float other = half == 0 ? __uint_as_float(s0[1]) : __uint_as_float(s1[0]);
        y[o] = own + other + s->b3[o];
    }
}

}  // namespace mpm
