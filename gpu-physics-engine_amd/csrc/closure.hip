// closure.hip — training and evaluation kernels of the learned constitutive closure (see closure.h).
//
// Training is plain minibatch Adam on the mean squared error of tau / mu, in FP32, with one kernel per
// layer pass.  Every reduction over samples runs in a fixed order (no atomics), so a gradient is
// bit-reproducible and can be checked against the FP64 restatement in the oracle.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/mpm_engine.h"
#include "closure.h"
#include "kernels.h"

namespace mpm {

__device__ __forceinline__ float sigmoid_f(float z) { return 1.f / (1.f + expf(-z)); }
__device__ __forceinline__ float silu_exact(float z) { return z * sigmoid_f(z); }
__device__ __forceinline__ float dsilu_exact(float z)
{
    const float s = sigmoid_f(z);
    return s * (1.f + z * (1.f - s));
}

// ---------------------------------------------------------------------------------------------- samples

// Features and targets (tau / mu from the analytic fixed-corotated / snow stress) of jelly and snow
// particles; valid[i] = 1 when the particle is inside the trained range.
__global__ void k_closure_samples(DevParams p, int n, const float* F, const float* J, const uint8_t* mat, float* X,
                                  float* T, int32_t* valid)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) { return; }
    const Mat2 f = {F[4 * i], F[4 * i + 1], F[4 * i + 2], F[4 * i + 3]};
    const int m = mat[i];
    const float jh = m == MPM_MAT_SNOW ? J[i] : 1.f;
    float x[4];
    const bool ok = m != MPM_MAT_FLUID && closure_features(f, jh, x);
    float mu = p.mu, lambda = p.lambda;
    if (m == MPM_MAT_SNOW) {
        const float h = expf(p.hardening * (1.f - J[i]));
        mu *= h;
        lambda *= h;
    }
    float t00 = 0.f, t01 = 0.f, t11 = 0.f;
    corotated_tau(f, mu, lambda, t00, t01, t11);
    const float inv_mu = 1.f / p.mu;
    for (int k = 0; k < 4; ++k) { X[4 * i + k] = ok ? x[k] : 0.f; }
    T[3 * i] = t00 * inv_mu;
    T[3 * i + 1] = t01 * inv_mu;
    T[3 * i + 2] = t11 * inv_mu;
    valid[i] = ok ? 1 : 0;
}

// ---------------------------------------------------------------------------------------------- training

// Minibatch indices: a SplitMix64-style hash of (seed, step, i) reduced to [0, n).
__device__ __forceinline__ uint32_t batch_index(uint64_t seed, int step, int i, int n)
{
    uint64_t z = seed + 0x9E3779B97F4A7C15ull * ((uint64_t)step * 0x100000000ull + (uint64_t)i + 1ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return (uint32_t)(z % (uint64_t)n);
}

__global__ void k_pick_batch(const float* X, const float* T, int n, int B, uint64_t seed, int step, int sequential,
                             float* Xb, float* Tb)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) { return; }
    const uint32_t s = sequential ? (uint32_t)i : batch_index(seed, step, i, n);
    for (int k = 0; k < 4; ++k) { Xb[4 * i + k] = X[4 * s + k]; }
    for (int k = 0; k < 3; ++k) { Tb[3 * i + k] = T[3 * s + k]; }
}

// Z1 = X W1^T + b1, H1 = silu(Z1)   (thread per sample x unit)
__global__ void k_fwd1(const float* w, const float* Xb, int B, float* Z1, float* H1)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * kHid) { return; }
    const int s = t / kHid, u = t % kHid;
    float z = w[kOffB1 + u];
    for (int k = 0; k < 4; ++k) { z = fmaf(w[kOffW1 + 4 * u + k], Xb[4 * s + k], z); }
    Z1[t] = z;
    H1[t] = silu_exact(z);
}
// Z2 = H1 W2^T + b2, H2 = silu(Z2)
__global__ void k_fwd2(const float* w, const float* H1, int B, float* Z2, float* H2)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * kHid) { return; }
    const int s = t / kHid, u = t % kHid;
    float z = w[kOffB2 + u];
    for (int k = 0; k < kHid; ++k) { z = fmaf(w[kOffW2 + kHid * u + k], H1[kHid * s + k], z); }
    Z2[t] = z;
    H2[t] = silu_exact(z);
}
// Y = H2 W3^T + b3; per-sample loss sum_j (y - t)^2 and dY = 2 (y - t) / B
__global__ void k_fwd3(const float* w, const float* H2, const float* Tb, int B, float* dY, float* loss)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B) { return; }
    float l = 0.f;
    for (int j = 0; j < 3; ++j) {
        float y = w[kOffB3 + j];
        for (int k = 0; k < kHid; ++k) { y = fmaf(w[kOffW3 + kHid * j + k], H2[kHid * s + k], y); }
        const float e = y - Tb[3 * s + j];
        l += e * e;
        dY[3 * s + j] = 2.f * e / (float)B;
    }
    loss[s] = l;
}
// dZ2 = (dY W3) * silu'(Z2)
__global__ void k_bwd2(const float* w, const float* dY, const float* Z2, int B, float* dZ2)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * kHid) { return; }
    const int s = t / kHid, u = t % kHid;
    float g = 0.f;
    for (int j = 0; j < 3; ++j) { g = fmaf(dY[3 * s + j], w[kOffW3 + kHid * j + u], g); }
    dZ2[t] = g * dsilu_exact(Z2[t]);
}
// dZ1 = (dZ2 W2) * silu'(Z1)
__global__ void k_bwd1(const float* w, const float* dZ2, const float* Z1, int B, float* dZ1)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * kHid) { return; }
    const int s = t / kHid, k = t % kHid;
    float g = 0.f;
    for (int u = 0; u < kHid; ++u) { g = fmaf(dZ2[kHid * s + u], w[kOffW2 + kHid * u + k], g); }
    dZ1[t] = g * dsilu_exact(Z1[t]);
}
// Parameter gradients, phase 1: work group b sums the contributions of samples [64 b, 64 b + 64) to every
// parameter (sample slice staged in LDS), writing partial[b][q].
constexpr int kGradSlice = 64;
__global__ __launch_bounds__(kThreads) void k_wgrad_partial(const float* Xb, const float* H1, const float* H2,
                                                            const float* dY, const float* dZ1, const float* dZ2,
                                                            int B, float* partial)
{
    __shared__ float sx[kGradSlice][4];
    __shared__ float sh1[kGradSlice][kHid];
    __shared__ float sh2[kGradSlice][kHid];
    __shared__ float sz1[kGradSlice][kHid];
    __shared__ float sz2[kGradSlice][kHid];
    __shared__ float sdy[kGradSlice][4];
    const int s0 = blockIdx.x * kGradSlice;
    const int ns = min(kGradSlice, B - s0);
    for (int t = threadIdx.x; t < kGradSlice * kHid; t += kThreads) {
        const int s = t / kHid, u = t % kHid;
        const bool ok = s < ns;
        const size_t g = (size_t)(s0 + s) * kHid + u;
        sh1[s][u] = ok ? H1[g] : 0.f;
        sh2[s][u] = ok ? H2[g] : 0.f;
        sz1[s][u] = ok ? dZ1[g] : 0.f;
        sz2[s][u] = ok ? dZ2[g] : 0.f;
    }
    for (int t = threadIdx.x; t < kGradSlice * 4; t += kThreads) {
        const int s = t / 4, k = t % 4;
        const bool ok = s < ns;
        sx[s][k] = ok ? Xb[(size_t)(s0 + s) * 4 + k] : 0.f;
        sdy[s][k] = (ok && k < 3) ? dY[(size_t)(s0 + s) * 3 + k] : 0.f;
    }
    __syncthreads();
    for (int q = threadIdx.x; q < kNumWeights; q += kThreads) {
        float g = 0.f;
        if (q < kOffB1) {
            const int u = q / 4, k = q % 4;
            for (int s = 0; s < kGradSlice; ++s) { g = fmaf(sz1[s][u], sx[s][k], g); }
        } else if (q < kOffW2) {
            const int u = q - kOffB1;
            for (int s = 0; s < kGradSlice; ++s) { g += sz1[s][u]; }
        } else if (q < kOffB2) {
            const int u = (q - kOffW2) / kHid, k = (q - kOffW2) % kHid;
            for (int s = 0; s < kGradSlice; ++s) { g = fmaf(sz2[s][u], sh1[s][k], g); }
        } else if (q < kOffW3) {
            const int u = q - kOffB2;
            for (int s = 0; s < kGradSlice; ++s) { g += sz2[s][u]; }
        } else if (q < kOffB3) {
            const int j = (q - kOffW3) / kHid, k = (q - kOffW3) % kHid;
            for (int s = 0; s < kGradSlice; ++s) { g = fmaf(sdy[s][j], sh2[s][k], g); }
        } else {
            const int j = q - kOffB3;
            for (int s = 0; s < kGradSlice; ++s) { g += sdy[s][j]; }
        }
        partial[(size_t)blockIdx.x * kNumWeights + q] = g;
    }
}
// Phase 2: G[q] = sum over work groups in order.
__global__ void k_wgrad_reduce(const float* partial, int nparts, float* G)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= kNumWeights) { return; }
    float g = 0.f;
    for (int b = 0; b < nparts; ++b) { g += partial[(size_t)b * kNumWeights + q]; }
    G[q] = g;
}
// Adam (Kingma & Ba 2015), bias-corrected, step t >= 1.
__global__ void k_adam(float* w, const float* G, float* m, float* v, float lr, float b1, float b2, float eps, int t)
{
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= kNumWeights) { return; }
    const float g = G[q];
    const float mq = b1 * m[q] + (1.f - b1) * g;
    const float vq = b2 * v[q] + (1.f - b2) * g * g;
    m[q] = mq;
    v[q] = vq;
    const float mh = mq / (1.f - powf(b1, (float)t));
    const float vh = vq / (1.f - powf(b2, (float)t));
    w[q] -= lr * mh / (sqrtf(vh) + eps);
}
// Per-sample losses summed in index order by one work group.
__global__ void k_sum(const float* a, int n, float* out)
{
    __shared__ float part[kThreads];
    float s = 0.f;
    for (int i = threadIdx.x; i < n; i += kThreads) { s += a[i]; }
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int k = 0; k < kThreads; ++k) { t += part[k]; }
        out[0] = t;
    }
}

// The network in MFMA operand layout (ClosurePacked): one thread per 16-byte A-operand entry or bias.
__global__ void k_pack_closure(const float* w, ClosurePacked* out)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = e % 64, m = lane % 32, h = lane / 32;
    bf16x8 v;
    constexpr float kC = kActScale;  // hidden pre-activations in index units of the SiLU table
    if (e < 2 * 64) {
        const int u = 32 * (e / 64) + m;
        for (int j = 0; j < 8; ++j) {
            const float f = kC * w[kOffW1 + 4 * u + j % 4];
            const __bf16 hi = (__bf16)f;
            v[j] = h ? (__bf16)(f - (float)hi) : hi;
        }
        out->a1[e] = v;
    } else if (e < 10 * 64) {
        const int blk = e / 64 - 2, mt = blk / 4, ks = blk % 4;
        for (int j = 0; j < 8; ++j) { v[j] = (__bf16)(kC * w[kOffW2 + kHid * (32 * mt + m) + chain_unit(ks, h, j)]); }
        out->a2[e - 2 * 64] = v;
    } else if (e < 10 * 64 + 32) {
        const int q = e - 10 * 64, ks = q / 8, hq = (q / 4) % 2, mq = q % 4;
        for (int j = 0; j < 8; ++j) { v[j] = (__bf16)(mq < 3 ? w[kOffW3 + kHid * mq + chain_unit(ks, hq, j)] : 0.f); }
        out->a3[q] = v;
    } else if (e < 10 * 64 + 32 + 128) {
        // bias of C register r of M-tile mt in half-wave hh
        const int b = e - (10 * 64 + 32), layer = b / 64, hh = (b / 32) % 2, mt = (b / 16) % 2, r = b % 16;
        const float bias = fmaf(kC, w[(layer ? kOffB2 : kOffB1) + 32 * mt + 8 * (r / 4) + 4 * hh + r % 4], kActShift);
        (layer ? out->b2p : out->b1p)[hh][16 * mt + r] = bias;
    } else if (e < 10 * 64 + 32 + 132) {
        const int o = e - (10 * 64 + 32) - 128;
        out->b3[o] = o < 3 ? w[kOffB3 + o] : 0.f;
    }
}

// SiLU segments in index units u = 16 z + 159.5 (see closure.h), computed once in FP64:
// segment i >= 1 covers z in [a, a + 1/16), a = -10 + i/16: the secant of silu over it, shifted by half its
// deviation at the midpoint (silu is within 1/256 of a parabola there, so the error is halved at the ends and
// the middle alike); segment 0 is the constant half of silu(-10 + 1/16), the middle of silu's range below.
static const float* act_table()
{
    static const std::vector<float> t = [] {
        std::vector<float> v(2 * kActSegs);
        const auto silu = [](double z) { return z / (1.0 + std::exp(-z)); };
        const double h = 1.0 / (double)kActScale, lo = (double)kActLo;
        for (int i = 0; i < kActSegs; ++i) {
            double slope_z = 0.0, icpt_z = 0.5 * silu(lo + h);
            if (i > 0) {
                const double a = lo + i * h, b = a + h, mid = a + 0.5 * h;
                slope_z = (silu(b) - silu(a)) / h;
                icpt_z = silu(a) - slope_z * a;
                icpt_z += 0.5 * (silu(mid) - (slope_z * mid + icpt_z));
            }
            // z = (u - shift) / scale: silu ~ slope_z z + icpt_z = (slope_z / scale) u + icpt_z - slope_z shift / scale
            v[2 * i] = (float)(slope_z / kActScale);
            v[2 * i + 1] = (float)(icpt_z - slope_z * kActShift / kActScale);
        }
        return v;
    }();
    return t.data();
}

// Network output of n feature vectors through the wave-level BF16 MFMA path used by the fused kernels.
__global__ __launch_bounds__(kThreads) void k_closure_eval(ClosureDev c, int n, const float* X, float* Y)
{
    __shared__ ClosurePacked lds;
    closure_stage(c, &lds);
    __syncthreads();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int ic = i < n ? i : n - 1;
    float x[4] = {X[4 * ic], X[4 * ic + 1], X[4 * ic + 2], X[4 * ic + 3]};
    for (int k = 0; k < 4; ++k) { x[k] = isfinite(x[k]) ? x[k] : 0.f; }  // the table index needs finite inputs
    float y[3];
    closure_eval_wave(&lds, x, y);
    if (i < n) {
        for (int k = 0; k < 3; ++k) { Y[3 * i + k] = y[k]; }
    }
}

size_t closure_packed_bytes() { return sizeof(ClosurePacked); }

hipError_t launch_pack_closure(hipStream_t st, const float* w, void* out)
{
    ClosurePacked* d = reinterpret_cast<ClosurePacked*>(out);
    hipLaunchKernelGGL(k_pack_closure, dim3((10 * 64 + 32 + 132 + kThreads - 1) / kThreads), dim3(kThreads), 0, st, w,
                       d);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { return e; }
    // the table is a static host array, so the asynchronous copy may read it at any later time
    return hipMemcpyAsync(d->act, act_table(), sizeof(d->act), hipMemcpyHostToDevice, st);
}

}  // namespace mpm

using namespace mpm;

namespace {

thread_local std::string g_closure_msg;

struct DevBuf {
    void* p = nullptr;
    ~DevBuf()
    {
        if (p) { (void)hipFree(p); }
    }
    template <class T>
    T* as()
    {
        return reinterpret_cast<T*>(p);
    }
};

#define CL_TRY(call)                                                                  \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) {                                                       \
            g_closure_msg = std::string(#call) + ": " + hipGetErrorName(e_);          \
            return MPM_ERR_DEVICE;                                                    \
        }                                                                             \
    } while (0)

inline unsigned nblk(size_t n) { return (unsigned)((n + kThreads - 1) / kThreads); }

// One forward + backward over B samples already in Xb/Tb; fills G and the per-sample losses.
hipError_t forward_backward(hipStream_t st, const float* w, const float* Xb, const float* Tb, int B, float* Z1,
                            float* H1, float* Z2, float* H2, float* dY, float* dZ1, float* dZ2, float* loss, float* G,
                            float* partial)
{
    const unsigned bh = nblk((size_t)B * kHid);
    hipLaunchKernelGGL(k_fwd1, dim3(bh), dim3(kThreads), 0, st, w, Xb, B, Z1, H1);
    hipLaunchKernelGGL(k_fwd2, dim3(bh), dim3(kThreads), 0, st, w, H1, B, Z2, H2);
    hipLaunchKernelGGL(k_fwd3, dim3(nblk(B)), dim3(kThreads), 0, st, w, H2, Tb, B, dY, loss);
    hipLaunchKernelGGL(k_bwd2, dim3(bh), dim3(kThreads), 0, st, w, dY, Z2, B, dZ2);
    hipLaunchKernelGGL(k_bwd1, dim3(bh), dim3(kThreads), 0, st, w, dZ2, Z1, B, dZ1);
    const int nparts = (B + kGradSlice - 1) / kGradSlice;
    hipLaunchKernelGGL(k_wgrad_partial, dim3(nparts), dim3(kThreads), 0, st, Xb, H1, H2, dY, dZ1, dZ2, B, partial);
    hipLaunchKernelGGL(k_wgrad_reduce, dim3(nblk(kNumWeights)), dim3(kThreads), 0, st, partial, nparts, G);
    return hipGetLastError();
}

struct TrainScratch {
    DevBuf buf;
    float *Xb, *Tb, *Z1, *H1, *Z2, *H2, *dY, *dZ1, *dZ2, *loss, *G, *sum, *partial;
    hipError_t alloc(int B)
    {
        const size_t hb = (size_t)B * kHid;
        const size_t nparts = (size_t)(B + kGradSlice - 1) / kGradSlice;
        const size_t total =
            (size_t)B * 4 + (size_t)B * 3 * 2 + hb * 6 + (size_t)B + kNumWeights + 1 + nparts * kNumWeights;
        hipError_t e = hipMalloc(&buf.p, total * sizeof(float));
        if (e != hipSuccess) { return e; }
        float* q = buf.as<float>();
        Xb = q;
        q += (size_t)B * 4;
        Tb = q;
        q += (size_t)B * 3;
        dY = q;
        q += (size_t)B * 3;
        Z1 = q;
        q += hb;
        H1 = q;
        q += hb;
        Z2 = q;
        q += hb;
        H2 = q;
        q += hb;
        dZ1 = q;
        q += hb;
        dZ2 = q;
        q += hb;
        loss = q;
        q += B;
        G = q;
        q += kNumWeights;
        sum = q;
        q += 1;
        partial = q;
        return hipSuccess;
    }
};

}  // namespace

extern "C" {

const char* mpm_closure_message(void) { return g_closure_msg.c_str(); }

int32_t mpm_closure_samples(const MpmParams* params, int32_t n, const float* F, const float* J,
                            const uint8_t* material, float* features, float* targets, int32_t* valid)
{
    if (!params || n < 0 || (n > 0 && (!F || !J || !material || !features || !targets || !valid))) {
        return MPM_ERR_INVALID_ARG;
    }
    if (n == 0) { return MPM_OK; }
    DevParams d{};
    d.mu = params->mu;
    d.lambda = params->lambda;
    d.hardening = params->hardening;
    const size_t N = (size_t)n;
    DevBuf b;
    CL_TRY(hipMalloc(&b.p, N * (4 + 1 + 4 + 3 + 1) * sizeof(float)));
    float* dF = b.as<float>();
    float* dJ = dF + 4 * N;
    float* dX = dJ + N;
    float* dT = dX + 4 * N;
    int32_t* dV = reinterpret_cast<int32_t*>(dT + 3 * N);
    DevBuf bm;
    CL_TRY(hipMalloc(&bm.p, N));
    CL_TRY(hipMemcpy(dF, F, 4 * N * sizeof(float), hipMemcpyHostToDevice));
    CL_TRY(hipMemcpy(dJ, J, N * sizeof(float), hipMemcpyHostToDevice));
    CL_TRY(hipMemcpy(bm.p, material, N, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_closure_samples, dim3(nblk(N)), dim3(kThreads), 0, 0, d, n, dF, dJ, bm.as<uint8_t>(), dX, dT,
                       dV);
    CL_TRY(hipGetLastError());
    CL_TRY(hipMemcpy(features, dX, 4 * N * sizeof(float), hipMemcpyDeviceToHost));
    CL_TRY(hipMemcpy(targets, dT, 3 * N * sizeof(float), hipMemcpyDeviceToHost));
    CL_TRY(hipMemcpy(valid, dV, N * sizeof(int32_t), hipMemcpyDeviceToHost));
    return MPM_OK;
}

int32_t mpm_closure_grad(int32_t n, const float* features, const float* targets, const float* weights, float* grad,
                         float* loss)
{
    if (n < 1 || !features || !targets || !weights || !grad || !loss) { return MPM_ERR_INVALID_ARG; }
    TrainScratch s;
    CL_TRY(s.alloc(n));
    DevBuf wb;
    CL_TRY(hipMalloc(&wb.p, kNumWeights * sizeof(float)));
    CL_TRY(hipMemcpy(wb.p, weights, kNumWeights * sizeof(float), hipMemcpyHostToDevice));
    CL_TRY(hipMemcpy(s.Xb, features, (size_t)n * 4 * sizeof(float), hipMemcpyHostToDevice));
    CL_TRY(hipMemcpy(s.Tb, targets, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice));
    CL_TRY(forward_backward(0, wb.as<float>(), s.Xb, s.Tb, n, s.Z1, s.H1, s.Z2, s.H2, s.dY, s.dZ1, s.dZ2, s.loss, s.G, s.partial));
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(kThreads), 0, 0, s.loss, n, s.sum);
    CL_TRY(hipGetLastError());
    CL_TRY(hipMemcpy(grad, s.G, kNumWeights * sizeof(float), hipMemcpyDeviceToHost));
    float tot = 0.f;
    CL_TRY(hipMemcpy(&tot, s.sum, sizeof(float), hipMemcpyDeviceToHost));
    *loss = tot / (float)n;
    return MPM_OK;
}

int32_t mpm_closure_train(int32_t n, const float* features, const float* targets, int32_t steps, int32_t batch,
                          float lr, uint64_t seed, float* weights, float* final_loss)
{
    if (n < 1 || !features || !targets || steps < 0 || batch < 1 || !(lr > 0.f) || !weights) {
        return MPM_ERR_INVALID_ARG;
    }
    TrainScratch s;
    CL_TRY(s.alloc(batch));
    DevBuf data, state;
    CL_TRY(hipMalloc(&data.p, (size_t)n * 7 * sizeof(float)));
    CL_TRY(hipMalloc(&state.p, 3 * kNumWeights * sizeof(float)));
    float* dX = data.as<float>();
    float* dT = dX + (size_t)n * 4;
    float* w = state.as<float>();
    float* m = w + kNumWeights;
    float* v = m + kNumWeights;
    CL_TRY(hipMemcpy(dX, features, (size_t)n * 4 * sizeof(float), hipMemcpyHostToDevice));
    CL_TRY(hipMemcpy(dT, targets, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice));
    CL_TRY(hipMemcpy(w, weights, kNumWeights * sizeof(float), hipMemcpyHostToDevice));
    CL_TRY(hipMemset(m, 0, 2 * kNumWeights * sizeof(float)));
    hipStream_t st = 0;
    for (int t = 1; t <= steps; ++t) {
        hipLaunchKernelGGL(k_pick_batch, dim3(nblk(batch)), dim3(kThreads), 0, st, dX, dT, n, batch, seed, t, 0, s.Xb,
                           s.Tb);
        CL_TRY(forward_backward(st, w, s.Xb, s.Tb, batch, s.Z1, s.H1, s.Z2, s.H2, s.dY, s.dZ1, s.dZ2, s.loss, s.G, s.partial));
        hipLaunchKernelGGL(k_adam, dim3(nblk(kNumWeights)), dim3(kThreads), 0, st, w, s.G, m, v, lr, 0.9f, 0.999f, 1e-8f,
                           t);
        CL_TRY(hipGetLastError());
    }
    if (final_loss) {
        // loss of the final weights on one more minibatch
        hipLaunchKernelGGL(k_pick_batch, dim3(nblk(batch)), dim3(kThreads), 0, st, dX, dT, n, batch, seed, steps + 1, 0,
                           s.Xb, s.Tb);
        CL_TRY(forward_backward(st, w, s.Xb, s.Tb, batch, s.Z1, s.H1, s.Z2, s.H2, s.dY, s.dZ1, s.dZ2, s.loss, s.G, s.partial));
        hipLaunchKernelGGL(k_sum, dim3(1), dim3(kThreads), 0, st, s.loss, batch, s.sum);
        float tot = 0.f;
        CL_TRY(hipMemcpy(&tot, s.sum, sizeof(float), hipMemcpyDeviceToHost));
        *final_loss = tot / (float)batch;
    }
    CL_TRY(hipMemcpy(weights, w, kNumWeights * sizeof(float), hipMemcpyDeviceToHost));
    return MPM_OK;
}

int32_t mpm_closure_act_table(float* slope_icpt)
{
    if (!slope_icpt) { return MPM_ERR_INVALID_ARG; }
    std::memcpy(slope_icpt, act_table(), 2 * kActSegs * sizeof(float));
    return MPM_OK;
}

int32_t mpm_closure_eval(int32_t n, const float* features, const float* weights, float* out)
{
    if (n < 1 || !features || !weights || !out) { return MPM_ERR_INVALID_ARG; }
    DevBuf b, packed;
    CL_TRY(hipMalloc(&b.p, ((size_t)n * 7 + kNumWeights) * sizeof(float)));
    CL_TRY(hipMalloc(&packed.p, sizeof(ClosurePacked)));
    float* dX = b.as<float>();
    float* dY = dX + (size_t)n * 4;
    float* dW = dY + (size_t)n * 3;
    CL_TRY(hipMemcpy(dX, features, (size_t)n * 4 * sizeof(float), hipMemcpyHostToDevice));
    CL_TRY(hipMemcpy(dW, weights, kNumWeights * sizeof(float), hipMemcpyHostToDevice));
    CL_TRY(launch_pack_closure(0, dW, packed.p));
    ClosureDev c{dW, packed.p};
    hipLaunchKernelGGL(k_closure_eval, dim3(nblk(n)), dim3(kThreads), 0, 0, c, n, dX, dY);
    CL_TRY(hipGetLastError());
    CL_TRY(hipMemcpy(out, dY, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost));
    return MPM_OK;
}

}  // extern "C"
