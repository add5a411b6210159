// mpm_device.h — device-side data layout and per-particle physics of the gfx950 MLS-MPM kernels.
//
// Everything here is __device__ inline and shared by the reference-structured, tiled and fused kernels,
// so the three execution modes compute identical per-particle arithmetic and differ only in where the
// grid values live (HBM or an LDS tile) and in the order of the floating-point atomic sums.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mpm_engine.h"

namespace mpm {

// ---- launch constants ----
constexpr int kThreads = 256;        // work-group size of every particle / node kernel (4 wave64)
constexpr int kRadixBits = 8;        // digit width of the binning sort
constexpr int kRadixBins = 1 << kRadixBits;
constexpr int kRadixItems = 16;      // keys per thread per sort tile
constexpr int kRadixTile = kThreads * kRadixItems;
constexpr int kMargin = 2;           // nodes of drift a binned particle may take before leaving its tile
constexpr int kAccTiles = kThreads / 64;  // LDS accumulation tiles per work group (one per wave)
constexpr int kMaxScenes = 4096;

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;

// Device copy of the learned closure's parameters (see closure.h).
struct ClosureDev {
    const float* w;       // float[4675]
    const void* packed;   // the network in MFMA operand layout (ClosurePacked)
};

// Kernel-argument copy of MpmParams plus derived quantities (passed by value, lives in SGPRs/kernarg).
struct DevParams {
    int W, H, halo, wall;
    int G;              // nodes per scene
    int num_scenes;
    int nbx, nby, bps;  // bins per scene along x, y and in total
    int bin_x, bin_y;   // cells per bin
    int node_bits;      // log2(bin_x * bin_y): bits of a stencil base node within its bin
    float dx, inv_dx, sim_w, sim_h, dt;
    float gdt_x, gdt_y;  // gravity * dt, folded once like the reference's constexpr kGravity * dt (MPM.cu:154)
    float mass, vol;
    float tait_k, tait_gamma, p_min, p_max;
    float mu, lambda, theta_c, theta_s, hardening;
    float j_min, j_max;
    int closure;
    ClosureDev cl;
};

// SoA particle store.  Every component array holds `capacity` entries, 256-byte aligned.
struct Store {
    float* x;
    float* y;
    float* vx;
    float* vy;
    float* c00;
    float* c01;
    float* c10;
    float* c11;
    float* f00;
    float* f01;
    float* f10;
    float* f11;
    float* J;       // tracked J (fluid, jelly) or plastic Jp (snow)
    uint8_t* mat;
    int32_t* id;    // caller's particle index
    uint32_t* key;  // bin key of the last binning (scene-major), carries the scene of the particle
};

// Grid of all scenes: scene s, node (x, y) at s*G + y*W + x (reference node index, MPM.cu:122).
struct Grid {
    float* mom_x;   // P2G accumulators this launch scatters into (zero when it starts)
    float* mom_y;
    float* mass;
    float2* vel;    // grid-update output read by G2P
    float* vmass;   // mass of the last grid update (readback)
    // Fused substep only (null otherwise): the accumulators of the previous substep, from which the
    // launch computes node velocities itself (the grid update folded into the G2P gather), and the
    // accumulators it zeroes for the substep after next (nodes per plane: zero_n).
    const float* src_x;
    const float* src_y;
    const float* src_m;
    float* zero_x;
    float* zero_y;
    float* zero_m;
    size_t zero_n;
};

// Work item of the binned kernels: particles [begin, end) of bin `key` (all within one bin).
struct Item {
    int32_t key;
    int32_t begin;
    int32_t end;
    int32_t pad;
};

struct Mat2 {
    float a00, a01, a10, a11;
};

// ---- B-spline weights (MPM.cu:92-110) ----
// The reference's gridCoord x * inv_dx + halo (MPM.cu:84) is contracted to one fused multiply-add by
// nvcc; it is written as an explicit fmaf here so the base node (and the bin key) is reproducible.
// With h = g - 0.5 (exact for the grid coordinates here), base = floor(h) and f = g - base = fract(h) + 0.5, so
// the weights' arguments 1.5 - f = 1 - r, f - 1 = r - 0.5 and f - 0.5 = r come from r = fract(h) exactly
// (r is a multiple of ulp(h), so 1 - r is exact).  base is taken by truncation, which equals floor for
// h >= 0; for h < 0 (positions left of the domain, possible only for emitted particles before their first
// out-of-bounds handling) truncation and floor both give a base <= 0, and every caller clamps the base to
// [0, dim - 3], so the clamped base is the same.  f and the weights use r either way.
__device__ __forceinline__ void bspline(float x, float inv_dx, float halo, int& base, float& f, float w[3])
{
    const float h = fmaf(x, inv_dx, halo) - 0.5f;
    base = (int)h;
    const float r = __builtin_amdgcn_fractf(h);
    f = r + 0.5f;
    const float a = 1.f - r;
    const float b = r - 0.5f;
    w[0] = 0.5f * a * a;
    w[1] = 0.75f - b * b;
    w[2] = 0.5f * r * r;
}

// v clamped to [lo, hi], lo <= hi: one v_max_i32 and one v_min_i32 (the nested select form compiles to three)
__device__ __forceinline__ int clampi(int v, int lo, int hi) { return min(max(v, lo), hi); }

// ---- 2x2 helpers (det2D / matMul2D, cuda-internal.cuh:44-56) ----
__device__ __forceinline__ float det2(const Mat2& m) { return m.a00 * m.a11 - m.a01 * m.a10; }
__device__ __forceinline__ Mat2 mul2(const Mat2& A, const Mat2& B)
{
    return {A.a00 * B.a00 + A.a01 * B.a10, A.a00 * B.a01 + A.a01 * B.a11, A.a10 * B.a00 + A.a11 * B.a10,
            A.a10 * B.a01 + A.a11 * B.a11};
}

// Rotation factor of the polar decomposition F = R S (the rotation maximising tr(R^T F)):
// angle atan2(F10 - F01, F00 + F11), R = I when both vanish.  Returns cos, sin.
// The FP32 form runs only for 0.1 <= det F <= 10, where r2 = |F|^2 + 2 det F >= 4 det F >= 0.4: the
// hardware reciprocal square root (v_rsq_f32, 1 ulp) needs no denormal scaling there.
__device__ __forceinline__ void polar_cs_f32(const Mat2& F, float& c, float& s)
{
    const float x = F.a00 + F.a11;
    const float y = F.a10 - F.a01;
    const float r2 = x * x + y * y;
    if (r2 > 0.f) {
        const float inv = __builtin_amdgcn_rsqf(r2);
        c = x * inv;
        s = y * inv;
    } else {
        c = 1.f;
        s = 0.f;
    }
}
// Same in FP64, used only for near-degenerate F (det F < 0.1 or det F > 10).
__device__ __forceinline__ void polar_cs_f64(const Mat2& F, float& c, float& s)
{
    const double x = (double)F.a00 + (double)F.a11;
    const double y = (double)F.a10 - (double)F.a01;
    const double r = sqrt(x * x + y * y);
    if (r > 0.0) {
        c = (float)(x / r);
        s = (float)(y / r);
    } else {
        c = 1.f;
        s = 0.f;
    }
}
__device__ __forceinline__ void polar_cs(const Mat2& F, float& c, float& s)
{
    const float J = det2(F);
    if (J < 0.1f || J > 10.f) {
        polar_cs_f64(F, c, s);
    } else {
        polar_cs_f32(F, c, s);
    }
}

// Fixed-corotated Kirchhoff stress tau = 2 mu (F - R) F^T + lambda (J - 1) J I, J = det F
// (Stomakhin et al. 2012), with R = [[c, -s], [s, c]] the polar rotation of F.  (F - R) F^T = F F^T - R F^T
// and R F^T = R S R^T is symmetric, so tau is symmetric; returns tau00, tau01 (= tau10), tau11.
__device__ __forceinline__ void corotated_tau_r(const Mat2& F, float c, float s, float mu, float lambda, float& t00,
                                                float& t01, float& t11)
{
    const float J = det2(F);
    // D = F - R, R = [[c, -s], [s, c]]
    const float d00 = F.a00 - c, d01 = F.a01 + s, d10 = F.a10 - s, d11 = F.a11 - c;
    const float m00 = d00 * F.a00 + d01 * F.a01;
    const float m01 = d00 * F.a10 + d01 * F.a11;
    const float m11 = d10 * F.a10 + d11 * F.a11;
    const float vol = lambda * (J - 1.f) * J;
    const float two_mu = 2.f * mu;
    t00 = two_mu * m00 + vol;
    t01 = two_mu * m01;
    t11 = two_mu * m11 + vol;
}
__device__ __forceinline__ void corotated_tau(const Mat2& F, float mu, float lambda, float& t00, float& t01,
                                              float& t11)
{
    float c, s;
    polar_cs(F, c, s);
    corotated_tau_r(F, c, s, mu, lambda, t00, t01, t11);
}

// Snow return mapping (Stomakhin et al. 2013): clamp the singular values of the trial elastic F to
// [1 - theta_c, 1 + theta_s], move the removed volume into Jp, rebuild F with the clamped values.
// The 2x2 SVD in closed form: every 2x2 matrix is a scaled rotation plus a scaled reflection,
//   F = [[E, -H], [H, E]] + [[D, G], [G, -D]],  2E = F00 + F11, 2H = F10 - F01, 2D = F00 - F11, 2G = F10 + F01,
// and with Q = |(E, H)|, R = |(D, G)| its signed SVD F = U diag(s0, s1) V^T (U, V rotations) has s0 = Q + R,
// s1 = Q - R (negative iff det F < 0), U V^T = Rot, the rotation by (E, H) / Q (the polar rotation of F),
// and U diag(1, -1) V^T = Refl, the reflection [[c, s], [s, -c]] with (c, s) = (D, G) / R.  Hence
// U diag(n0, n1) V^T = (n0 + n1)/2 Rot + (n0 - n1)/2 Refl: neither U nor V is formed.  (rc, rs) is the polar
// rotation from polar_cs (FP64 for det F outside [0.1, 10]); the clamped singular values are positive, so it is
// also the polar rotation of the rebuilt F.  A non-finite F stays non-finite, as the clamp alone would not keep.
__device__ __forceinline__ void snow_project(const DevParams& p, Mat2& F, float& Jp, float& rc, float& rs)
{
    polar_cs(F, rc, rs);
    const float q = 0.5f * ((F.a00 + F.a11) * rc + (F.a10 - F.a01) * rs);  // Q = (E, H) . (rc, rs)
    const float d2 = F.a00 - F.a11, g2 = F.a10 + F.a01;
    const float r2 = d2 * d2 + g2 * g2;  // (2R)^2
    // below (2R)^2 = 4e-30 F is a scaled rotation to within 1e-15 and the reflection part is dropped; above it
    // r2 is a normal number and the hardware reciprocal square root (1 ulp) needs no denormal scaling
    float fc = 1.f, fs = 0.f, r = 0.f;
    if (r2 > 4e-30f) {
        const float inv = __builtin_amdgcn_rsqf(r2);
        fc = d2 * inv;
        fs = g2 * inv;
        r = 0.5f * r2 * inv;
    }
    const float s0 = q + r, s1 = q - r;
    const float lo = 1.f - p.theta_c, hi = 1.f + p.theta_s;
    const float n0 = fminf(fmaxf(s0, lo), hi);
    const float n1 = fminf(fmaxf(s1, lo), hi);
    Jp *= (s0 * s1) * __builtin_amdgcn_rcpf(n0 * n1);  // n0 n1 in [lo^2, hi^2], near 1
    const float kr = fmaf(0.f, q, 0.5f * (n0 + n1));  // 0 q: NaN or inf in F reaches the rebuilt F
    const float kf = 0.5f * (n0 - n1);
    F.a00 = kr * rc + kf * fc;
    F.a01 = kf * fs - kr * rs;
    F.a10 = kr * rs + kf * fs;
    F.a11 = kr * rc - kf * fc;
}

// Tait pressure of the tracked J (MPM.cu:68-69): p = k ((1/J)^gamma - 1) clamped to [p_min, p_max].  1/J is
// the hardware reciprocal (1 ulp; every G2P clamps J to [j_min, j_max], j_min > 0); gamma = 1 (linear bulk
// modulus) skips powf, whose result would be 1/J exactly.
__device__ __forceinline__ float tait_pressure(const DevParams& p, float J)
{
    const float inv_j = __builtin_amdgcn_rcpf(J);
    const float p_hat = p.tait_k * ((p.tait_gamma == 1.f ? inv_j : powf(inv_j, p.tait_gamma)) - 1.f);
    return fminf(fmaxf(p_hat, p.p_min), p.p_max);
}

// What the return mapping of a snow particle leaves for its stress: the polar rotation (c, s) of the projected F
// and the hardening factor h = exp(hardening (1 - Jp)) of its new Jp.
struct SnowHint {
    float c, s, h;
};
__device__ __forceinline__ float snow_hardening(const DevParams& p, float Jp) { return __expf(p.hardening * (1.f - Jp)); }

// Kirchhoff stress tau = (t00, t01 = t10, t11) of one particle: the stress of every P2G kernel and of the
// stress diagnostic, so both run the same instructions.
//   fluid: tau = J sigma = -J p I with the Tait pressure p (the reference's stress, MPM.cu:68-80, where
//          Q = m C + (-dt V0 J 4 inv_dx^2)(-p I) = m C - dt V0 4 inv_dx^2 tau);
//   jelly: fixed-corotated on F with the polar rotation of F;
//   snow:  the same with mu, lambda scaled by exp(hardening (1 - Jp)) (Stomakhin et al. 2013).
// snow: for a snow particle, the rotation and hardening factor when the caller holds them from the return
// mapping (g2p_update), else nullptr and they are computed here (the same expressions).
__device__ __forceinline__ void kirchhoff_tau(const DevParams& p, int mat, const Mat2& F, float J,
                                              const SnowHint* snow, float& t00, float& t01, float& t11)
{
    if (mat == MPM_MAT_FLUID) {
        const float t = -J * tait_pressure(p, J);
        t00 = t;
        t01 = 0.f;
        t11 = t;
        return;
    }
    float mu = p.mu, lambda = p.lambda;
    float c, s;
    if (mat == MPM_MAT_SNOW && snow) {
        c = snow->c;
        s = snow->s;
        mu *= snow->h;
        lambda *= snow->h;
    } else {
        if (mat == MPM_MAT_SNOW) {
            const float h = snow_hardening(p, J);
            mu *= h;
            lambda *= h;
        }
        polar_cs(F, c, s);
    }
    corotated_tau_r(F, c, s, mu, lambda, t00, t01, t11);
}

// Affine momentum matrix Q = m C - dt V0 4 inv_dx^2 tau of P2G (MPM.cu:71-80) with the stress of
// kirchhoff_tau.
__device__ __forceinline__ Mat2 affine_q(const DevParams& p, int mat, const Mat2& C, const Mat2& F, float J,
                                         const SnowHint* snow = nullptr)
{
    float t00, t01, t11;
    kirchhoff_tau(p, mat, F, J, snow, t00, t01, t11);
    const float D = p.dt * p.vol * 4.f * p.inv_dx * p.inv_dx;
    return {p.mass * C.a00 - D * t00, p.mass * C.a01 - D * t01, p.mass * C.a10 - D * t01, p.mass * C.a11 - D * t11};
}

// handleOOB (MPM.cu:267-301), one axis: a non-finite position or velocity component resets the position to
// the middle of the domain and the velocity to 0; then x < 0 sets x = 0, v = max(v, 0) and x > w sets x = w,
// v = min(v, 0) (w > 0, so at most one applies).  Written as two med3 clamps: v is finite there, and
// med3(v, 0, +inf) = max(v, 0), med3(v, -inf, 0) = min(v, 0), med3(x, 0, w) = the clamped x (up to the sign
// of a zero result).
__device__ __forceinline__ void oob_axis(float& x, float& v, float w)
{
    if (!isfinite(x) || !isfinite(v)) {
        x = 0.5f * w;
        v = 0.f;
    }
    const float lo = x < 0.f ? 0.f : -INFINITY;
    const float hi = x > w ? 0.f : INFINITY;
    v = __builtin_amdgcn_fmed3f(v, lo, hi);
    x = __builtin_amdgcn_fmed3f(x, 0.f, w);
}
__device__ __forceinline__ void handle_oob(const DevParams& p, float& px, float& py, float& vx, float& vy)
{
    oob_axis(px, vx, p.sim_w);
    oob_axis(py, vy, p.sim_h);
}

// Particle update after the gather (MPM.cu:252-265): advect, out-of-bounds, F <- (I + dt C) F, then
// J <- clamp(J det A) or, for snow, the return mapping on F and Jp.  C = 4 inv_dx^2 B (MPM.cu:255) comes from
// the caller's gather (formed before the OOB handling, as in the reference).  For snow, *snow (if given)
// receives the polar rotation of the new F and the hardening factor of the new Jp, for affine_q.
__device__ __forceinline__ void g2p_update(const DevParams& p, int mat, float& px, float& py, float& vx, float& vy,
                                           const Mat2& C, Mat2& F, float& J, SnowHint* snow = nullptr)
{
    px += vx * p.dt;
    py += vy * p.dt;
    handle_oob(p, px, py, vx, vy);
    const Mat2 A = {1.f + C.a00 * p.dt, C.a01 * p.dt, C.a10 * p.dt, 1.f + C.a11 * p.dt};
    F = mul2(A, F);
    if (mat == MPM_MAT_SNOW) {
        float rc, rs;
        snow_project(p, F, J, rc, rs);
        if (snow) { *snow = {rc, rs, snow_hardening(p, J)}; }
    } else {
        J *= det2(A);
        J = fminf(fmaxf(J, p.j_min), p.j_max);
    }
}

// Grid node update (MPM.cu:150-172): momentum / mass, gravity, outward-only walls, empty node -> 0.
__device__ __forceinline__ float2 node_velocity(const DevParams& p, float mx, float my, float m, int x, int y)
{
    float2 v = {0.f, 0.f};
    if (m > 0.f) {
        const float inv_m = 1.f / m;
        v.x = mx * inv_m + p.gdt_x;
        v.y = my * inv_m + p.gdt_y;
        if (x <= p.wall && v.x < 0.f) {
            v.x = 0.f;
        } else if (x >= p.W - 1 - p.wall && v.x > 0.f) {
            v.x = 0.f;
        }
        if (y <= p.wall && v.y < 0.f) {
            v.y = 0.f;
        } else if (y >= p.H - 1 - p.wall && v.y > 0.f) {
            v.y = 0.f;
        }
    }
    return v;
}

// Bin key of a position (scene-major) and the stencil base node within the bin (row-major, x fastest);
// base node clamped for memory safety.
__device__ __forceinline__ uint32_t bin_key(const DevParams& p, float x, float y, uint32_t scene, uint32_t& node)
{
    int bx = (int)floorf(fmaf(x, p.inv_dx, (float)p.halo) - 0.5f);
    int by = (int)floorf(fmaf(y, p.inv_dx, (float)p.halo) - 0.5f);
    bx = clampi(bx, 0, p.W - 3);
    by = clampi(by, 0, p.H - 3);
    node = (uint32_t)((by % p.bin_y) * p.bin_x + bx % p.bin_x);
    return scene * (uint32_t)p.bps + (uint32_t)((by / p.bin_y) * p.nbx + bx / p.bin_x);
}

}  // namespace mpm
