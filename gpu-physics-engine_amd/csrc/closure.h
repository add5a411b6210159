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

// Hidden activations.  The packed network computes every hidden pre-activation z in index units
// u = kActScale z + kActShift (W1, W2 scaled by 16, a power of two, so their BF16 parts are those of the
// unscaled weights; biases 16 b + 159.5).  SiLU is then piecewise linear over 256 segments of 1/16 in z:
// segment i = clamp(rint(u), 0, 255), which covers z in [-10 + i/16, -10 + (i + 1)/16), holds a line
// silu ~ slope_i u + icpt_i read from an LDS table.  The index is one v_cvt_pk_u8_f32 (round to nearest even,
// saturated to [0, 255]), so an activation costs that conversion, a shift, one LDS read and one FMA instead of
// the two transcendentals of z / (1 + e^-z).  Segments 1..255 cover z in [-10 + 1/16, 6) with the secant
// shifted by half its midpoint deviation (|error| <= 1.3e-4), segment 255 extends linearly above z = 6, and
// segment 0 is the constant silu(-10 + 1/16) / 2 for every z below (|error| <= 2.4e-4).  The lines are
// computed in FP64 on the host (mpm_closure_act_table).
constexpr int kActSegs = 256;
constexpr float kActScale = 16.f;
constexpr float kActLo = -10.f;                            // z at the start of segment 0
constexpr float kActShift = -kActLo * kActScale - 0.5f;    // 159.5: rint(u) = floor(16 (z + 10))

// atan2(y, x): atan on [0, 1] from the degree-9 odd polynomial of Abramowitz & Stegun 4.4.47 (|error| <=
// 1e-5 rad) plus octant reduction.  Training and inference use this same function, so the network's angle
// input is consistent.
__device__ __forceinline__ float fast_atan2(float y, float x)
{
    const float ax = fabsf(x), ay = fabsf(y);
    const float mx = fmaxf(ax, ay), mn = fminf(ax, ay);
    const float t = mx > 0.f ? mn * __builtin_amdgcn_rcpf(mx) : 0.f;
    const float t2 = t * t;
    float r = 0.0208351f;
    r = fmaf(r, t2, -0.0851330f);
    r = fmaf(r, t2, 0.1801410f);
    r = fmaf(r, t2, -0.3302995f);
    r = fmaf(r, t2, 0.9998660f);
    r *= t;
    if (ay > ax) { r = 1.57079632679f - r; }
    if (x < 0.f) { r = 3.14159265359f - r; }
    return y < 0.f ? -r : r;
}

// Closure features of one particle; returns false when outside the trained range (then x is unspecified).
// The SVD comes in closed form: with E = (F00 + F11) / 2, D = (F00 - F11) / 2, G = (F10 + F01) / 2 and
// H = (F10 - F01) / 2, F = U diag(s0, s1) V^T with s0 = q + r >= s1 = q - r for q = |(E, H)|, r = |(D, G)|,
// and U the rotation by (atan2(H, E) + atan2(G, D)) / 2.  tau = U diag U^T is unchanged by U -> -U, so only
// that angle mod pi matters: it is half the argument of (E + iH)(D + iG) = (ED - HG) + i(EG + HD), one atan2,
// and atan2 / 2 lies in (-pi/2, pi/2] already.  At s0 = s1 any U serves (there tau is isotropic).  In the
// trained range q >= 0.6 (q is at least the mean singular value), so the hardware square root needs no
// denormal scaling.
__device__ __forceinline__ bool closure_features(const Mat2& F, float jh, float x[4])
{
    const float E = 0.5f * (F.a00 + F.a11), D = 0.5f * (F.a00 - F.a11);
    const float G = 0.5f * (F.a10 + F.a01), H = 0.5f * (F.a10 - F.a01);
    const float q = __builtin_amdgcn_sqrtf(fmaf(E, E, H * H));
    const float r = __builtin_amdgcn_sqrtf(fmaf(D, D, G * G));
    const float s0 = q + r, s1 = q - r;
    const float th = 0.5f * fast_atan2(fmaf(E, G, H * D), fmaf(E, D, -H * G));
    x[0] = (s0 - 1.1f) * 2.f;
    x[1] = (s1 - 1.1f) * 2.f;
    x[2] = th * 0.636619772368f;
    x[3] = (jh - 1.f) * (1.f / 0.3f);
    return s0 >= kSigLo && s0 <= kSigHi && s1 >= kSigLo && s1 <= kSigHi && jh >= kJhLo && jh <= kJhHi;
}

// Hidden unit fed as K index kk = 8 h + j of K-step ks: the C register r = 8 (ks % 2) + j of M-tile ks / 2
// of the previous layer, whose unit is 32 (ks/2) + 8 (r/4) + 4 h + r%4.  With this order a layer's FP32
// output (C layout) is the next layer's B operand lane by lane, so layers chain without lane exchanges.
__host__ __device__ constexpr int chain_unit(int ks, int h, int j)
{
    return 32 * (ks / 2) + 16 * (ks % 2) + 8 * (j / 4) + 4 * h + (j % 4);
}

// Network operands in MFMA layout (v_mfma_f32_32x32x16_bf16: A lane l holds A[l%32][8(l/32)+j], B lane l
// holds B[8(l/32)+j][l%32], C lane l holds C[8(r/4)+4(l/32)+r%4][l%32]; checked on the device).
struct ClosurePacked {
    float2 act[kActSegs];  // SiLU segments (slope, intercept) in index units u
    // with c = kActScale, W1s = c W1 split as W1s = W1hi + W1lo (BF16 parts):
    bf16x8 a1[2 * 64];  // layer 1, M-tile mt: A[u][8h+j] = (h ? W1lo : W1hi)[u][j % 4], u = 32 mt + l%32
    bf16x8 a2[8 * 64];  // layer 2, (mt, ks): A[m][8h+j] = W2[32 mt + l%32][chain_unit(ks, h, j)]
    bf16x8 a3[4 * 8];   // layer 3, ks: [8 ks + 4 h + m] = A[m][8h+j] = W3[m][chain_unit(ks, h, j)] for the
                        // three output rows m < 3, zero for m = 3; rows 3..31 of A are zero and read entry m = 3
    float b1p[2][32];   // [h][16 mt + r] = c b1[32 mt + 8 (r/4) + 4 h + r%4] + kActShift
    float b2p[2][32];   // likewise c b2 + kActShift
    float b3[4];
};

__device__ __forceinline__ void closure_stage(const ClosureDev& c, ClosurePacked* s)
{
    const uint4* src = reinterpret_cast<const uint4*>(c.packed);
    uint4* dst = reinterpret_cast<uint4*>(s);
    constexpr int n16 = (int)(sizeof(ClosurePacked) / 16);
    for (int i = threadIdx.x; i < n16; i += blockDim.x) { dst[i] = src[i]; }
}

// Keeps the LDS operand loads of one network stage next to their MFMAs: without it the compiler hoists
// all 14 A operands and 64 biases out of the particle loop and the fused kernels lose occupancy.
__device__ __forceinline__ void closure_stage_fence() { __asm__ __volatile__("" ::: "memory"); }

// Accumulator of M-tile mt initialised with the layer's biases (C layout, this half-wave's rows).
__device__ __forceinline__ floatx16 bias_acc(const float (*bias)[32], int half, int mt)
{
    const float4* b = reinterpret_cast<const float4*>(&bias[half][16 * mt]);
    floatx16 acc;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const float4 v = b[q];
        acc[4 * q] = v.x;
        acc[4 * q + 1] = v.y;
        acc[4 * q + 2] = v.z;
        acc[4 * q + 3] = v.w;
    }
    return acc;
}

// Layer output -> next B operand: SiLU and BF16, for K-steps 2 mt and 2 mt + 1.  All 16 table reads are
// issued before the first FMA so their LDS latencies overlap.  The saturating byte conversion keeps every
// index in [0, 255] (u is finite in every lane: the features of every lane are).
__device__ __forceinline__ void chain_operands(const floatx16& acc, const float2* act, bf16x8& lo, bf16x8& hi)
{
    float2 c[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) { c[j] = act[__builtin_amdgcn_cvt_pk_u8_f32(acc[j], 0u, 0u)]; }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        lo[j] = (__bf16)fmaf(c[j].x, acc[j], c[j].y);
        hi[j] = (__bf16)fmaf(c[8 + j].x, acc[8 + j], c[8 + j].y);
    }
}

// Network output (tau / mu) of the particle of every lane.  Must be called by all 64 lanes
// (wave-uniform control flow); x holds the lane's own features.  Particles are the N dimension of the
// MFMAs: N-tile nt holds particles 32 nt + l%32, evaluated one N-tile after the other.  The N-tile loop is
// unrolled so that nt is a constant: a runtime nt costs selects on every operand and result it picks.
__device__ __forceinline__ void closure_eval_wave(const ClosurePacked* s, const float x[4], float y[3])
{
    const int lane = threadIdx.x & 63;
    const int half = lane >> 5;
    // layer 1 B operand: (x_hi, x_lo) of particle 32 nt + l%32 in both halves, so that
    // A1 B1 = W1hi (x_hi + x_lo) + W1lo (x_hi + x_lo).  Each lane splits its own features once; the two
    // half-wave exchanges of the packed words then give the operands of both N-tiles.
    bf16x8 own;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const __bf16 fh = (__bf16)x[k];
        own[k] = fh;
        own[4 + k] = (__bf16)(x[k] - (float)fh);
    }
    typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 ow = __builtin_bit_cast(u32x4, own);
    u32x4 op[2];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const auto sw = __builtin_amdgcn_permlane32_swap(ow[q], ow[q], false, false);
        op[0][q] = sw[0];
        op[1][q] = sw[1];
    }
    float res[2][3];
#pragma unroll
    for (int nt = 0; nt < 2; ++nt) {
        const bf16x8 b1 = __builtin_bit_cast(bf16x8, op[nt]);
        bf16x8 b2[4];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            closure_stage_fence();
            floatx16 acc = bias_acc(s->b1p, half, mt);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(s->a1[mt * 64 + lane], b1, acc, 0, 0, 0);
            chain_operands(acc, s->act, b2[2 * mt], b2[2 * mt + 1]);
        }
        bf16x8 b3[4];
#pragma unroll
        for (int mt = 0; mt < 2; ++mt) {
            closure_stage_fence();
            floatx16 acc = bias_acc(s->b2p, half, mt);
#pragma unroll
            for (int ks = 0; ks < 4; ++ks) {
                acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(s->a2[(mt * 4 + ks) * 64 + lane], b2[ks], acc, 0, 0, 0);
            }
            chain_operands(acc, s->act, b3[2 * mt], b3[2 * mt + 1]);
        }
        closure_stage_fence();
        floatx16 out = {};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
            out = __builtin_amdgcn_mfma_f32_32x32x16_bf16(s->a3[ks * 8 + half * 4 + min(lane & 31, 3)], b3[ks], out,
                                                          0, 0, 0);
        }
        // rows 0..2 (the three outputs) sit in registers 0..2 of the lower half-wave
        res[nt][0] = out[0];
        res[nt][1] = out[1];
        res[nt][2] = out[2];
    }
    closure_stage_fence();
    // lower half: own particle l is in N-tile 0; upper half: own particle l is N-tile 1 of lane l - 32
#pragma unroll
    for (int o = 0; o < 3; ++o) {
        const auto sw = __builtin_amdgcn_permlane32_swap(__float_as_uint(res[1][o]), __float_as_uint(res[1][o]), false,
                                                         false);
        y[o] = (half == 0 ? res[0][o] : __uint_as_float(sw[0])) + s->b3[o];
    }
}

}  // namespace mpm

