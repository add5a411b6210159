// kernels.hip — gfx950 kernels of the MLS-MPM substep.
//
// Reference kernels (libs/cuda-engine/source/MPM.cu) and their counterparts here:
//   resetGridKernel   :46-51   -> folded into k_grid_update (it zeroes the accumulators it consumed)
//   P2GKernel         :53-129  -> k_p2g_reference (same structure: one particle per lane, 27 HBM atomics)
//                                 k_p2g_tiled     (binned particles, LDS tile, one HBM atomic per tile node)
//   gridKernel        :131-177 -> k_grid_update
//   G2PKernel         :178-301 -> k_g2p_reference / k_g2p_tiled (LDS-staged velocity tile)
//   (P2G of s+1 fused with G2P of s, no reference counterpart) -> k_g2p2g
//   MPMOutputKernel   :302-322 -> k_output
//   MPMEmitterKernel  :380-438 -> k_emitter (own Philox4x32-10 + Box-Muller, wave64 ballot)
// plus the binning that the reference MPM does not have: a stable radix sort by (bin, stencil base node)
// and a deterministic layered order inside each bin (the reference's only sort, PBF.cu:738-834, ranks with
// atomicAdd and is order-nondeterministic).
//
// The plain fused kernel k_g2p2g is compiled in a translation unit of its own (g2p2g.hip includes this file
// with MPM_PLAIN_FUSED_TU defined) so that it alone can take LLVM's max-ilp machine scheduler; everything
// that must exist once (kernels with external linkage, launchers, device variables) is guarded by
// MPM_MAIN_TU, the shared device code and the kernel templates are seen by both.
#include "kernels.h"
#include "closure.h"

#include <hip/hip_runtime.h>

#ifndef MPM_PLAIN_FUSED_TU
#define MPM_MAIN_TU 1
#endif

namespace mpm {

// ===================================================================================== small utilities

__device__ __forceinline__ int scene_of_key(const DevParams& p, uint32_t key)
{
    return p.num_scenes > 1 ? (int)(key / (uint32_t)p.bps) : 0;
}

__device__ __forceinline__ void atomic_add_global(float* a, float v) { unsafeAtomicAdd(a, v); }

#ifdef MPM_MAIN_TU

__global__ void k_fill_f32(float* p, size_t n, float v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) { p[i] = v; }
}
__global__ void k_fill_i32(int32_t* p, size_t n, int32_t v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) { p[i] = v; }
}

static unsigned fill_blocks(size_t n)
{
    size_t b = (n + kThreads - 1) / kThreads;
    if (b > 65536) { b = 65536; }
    return (unsigned)(b ? b : 1);
}

hipError_t launch_fill_f32(hipStream_t st, float* p, size_t n, float v)
{
    if (n == 0) { return hipSuccess; }
    hipLaunchKernelGGL(k_fill_f32, dim3(fill_blocks(n)), dim3(kThreads), 0, st, p, n, v);
    return hipGetLastError();
}
hipError_t launch_fill_i32(hipStream_t st, int32_t* p, size_t n, int32_t v)
{
    if (n == 0) { return hipSuccess; }
    hipLaunchKernelGGL(k_fill_i32, dim3(fill_blocks(n)), dim3(kThreads), 0, st, p, n, v);
    return hipGetLastError();
}

// ===================================================================================== grid update

// gridKernel (MPM.cu:131-177) for every node of every scene, plus the reset of the accumulators it
// consumed (resetGridKernel, MPM.cu:46-51), so the next P2G starts from zero without a separate pass; in
// the fused substep it also zeroes the accumulator set of the substep before (g.zero_*).
// Each lane handles 4 consecutive nodes (float4 loads of the accumulator planes).
__global__ __launch_bounds__(kThreads) void k_grid_update(DevParams p, Grid g, size_t total)
{
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    const size_t i0 = q * 4;
    if (i0 >= total) { return; }
    if (i0 + 4 <= total) {
        float4 mx = *reinterpret_cast<const float4*>(g.mom_x + i0);
        float4 my = *reinterpret_cast<const float4*>(g.mom_y + i0);
        float4 m = *reinterpret_cast<const float4*>(g.mass + i0);
        const float amx[4] = {mx.x, mx.y, mx.z, mx.w};
        const float amy[4] = {my.x, my.y, my.z, my.w};
        const float am[4] = {m.x, m.y, m.z, m.w};
        float2 out[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const size_t idx = i0 + k;
            const int local = (int)(idx % (size_t)p.G);
            const int y = local / p.W;
            const int x = local - y * p.W;
            out[k] = node_velocity(p, amx[k], amy[k], am[k], x, y);
        }
        float4* vel4 = reinterpret_cast<float4*>(g.vel + i0);
        vel4[0] = make_float4(out[0].x, out[0].y, out[1].x, out[1].y);
        vel4[1] = make_float4(out[2].x, out[2].y, out[3].x, out[3].y);
        *reinterpret_cast<float4*>(g.vmass + i0) = m;
        const float4 z = make_float4(0.f, 0.f, 0.f, 0.f);
        *reinterpret_cast<float4*>(g.mom_x + i0) = z;
        *reinterpret_cast<float4*>(g.mom_y + i0) = z;
        *reinterpret_cast<float4*>(g.mass + i0) = z;
        if (g.zero_x) {
            *reinterpret_cast<float4*>(g.zero_x + i0) = z;
            *reinterpret_cast<float4*>(g.zero_y + i0) = z;
            *reinterpret_cast<float4*>(g.zero_m + i0) = z;
        }
    } else {
        for (size_t idx = i0; idx < total; ++idx) {
            const int local = (int)(idx % (size_t)p.G);
            const int y = local / p.W;
            const int x = local - y * p.W;
            g.vel[idx] = node_velocity(p, g.mom_x[idx], g.mom_y[idx], g.mass[idx], x, y);
            g.vmass[idx] = g.mass[idx];
            g.mom_x[idx] = 0.f;
            g.mom_y[idx] = 0.f;
            g.mass[idx] = 0.f;
            if (g.zero_x) {
                g.zero_x[idx] = 0.f;
                g.zero_y[idx] = 0.f;
                g.zero_m[idx] = 0.f;
            }
        }
    }
}

hipError_t launch_grid_update(hipStream_t st, const DevParams& p, const Grid& g)
{
    const size_t total = (size_t)p.G * (size_t)p.num_scenes;
    const size_t lanes = (total + 3) / 4;
    hipLaunchKernelGGL(k_grid_update, dim3((unsigned)((lanes + kThreads - 1) / kThreads)), dim3(kThreads), 0, st, p,
                       g, total);
    return hipGetLastError();
}

#endif  // MPM_MAIN_TU

// ===================================================================================== per-particle helpers

__device__ __forceinline__ Mat2 load_F(const Store& s, int i) { return {s.f00[i], s.f01[i], s.f10[i], s.f11[i]}; }
__device__ __forceinline__ Mat2 load_C(const Store& s, int i) { return {s.c00[i], s.c01[i], s.c10[i], s.c11[i]}; }
__device__ __forceinline__ void store_F(const Store& s, int i, const Mat2& F)
{
    s.f00[i] = F.a00;
    s.f01[i] = F.a01;
    s.f10[i] = F.a10;
    s.f11[i] = F.a11;
}
__device__ __forceinline__ void store_C(const Store& s, int i, const Mat2& C)
{
    s.c00[i] = C.a00;
    s.c01[i] = C.a01;
    s.c10[i] = C.a10;
    s.c11[i] = C.a11;
}

// SoA element access through the store's base pointer and a 32-bit byte offset: the loads and stores
// use the SGPR-base + VGPR-offset form, with no 64-bit address arithmetic per component.
template <class T>
__device__ __forceinline__ T ldo(const T* base, uint32_t off)
{
    return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}
template <class T>
__device__ __forceinline__ void sto(T* base, uint32_t off, T v)
{
    *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + off) = v;
}
__device__ __forceinline__ Mat2 ld_F(const Store& s, uint32_t o)
{
    return {ldo(s.f00, o), ldo(s.f01, o), ldo(s.f10, o), ldo(s.f11, o)};
}
__device__ __forceinline__ Mat2 ld_C(const Store& s, uint32_t o)
{
    return {ldo(s.c00, o), ldo(s.c01, o), ldo(s.c10, o), ldo(s.c11, o)};
}
__device__ __forceinline__ void st_F(const Store& s, uint32_t o, const Mat2& F)
{
    sto(s.f00, o, F.a00);
    sto(s.f01, o, F.a01);
    sto(s.f10, o, F.a10);
    sto(s.f11, o, F.a11);
}
__device__ __forceinline__ void st_C(const Store& s, uint32_t o, const Mat2& C)
{
    sto(s.c00, o, C.a00);
    sto(s.c01, o, C.a01);
    sto(s.c10, o, C.a10);
    sto(s.c11, o, C.a11);
}

#ifdef MPM_MAIN_TU
// ===================================================================================== reference structure

// P2GKernel + particleToGrid (MPM.cu:53-129): one particle per lane, 27 float atomics into HBM.
__global__ __launch_bounds__(kThreads) void k_p2g_reference(DevParams p, Store s, const int32_t* count, Grid g)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count[0]) { return; }
    const int scene = scene_of_key(p, s.key[i]);
    const size_t gb = (size_t)scene * (size_t)p.G;
    const int mat = s.mat[i];
    const Mat2 Q = affine_q(p, mat, load_C(s, i), load_F(s, i), s.J[i]);
    const float vx = s.vx[i], vy = s.vy[i];
    int bx, by;
    float fx, fy, wx[3], wy[3];
    bspline(s.x[i], p.inv_dx, (float)p.halo, bx, fx, wx);
    bspline(s.y[i], p.inv_dx, (float)p.halo, by, fy, wy);
    bx = clampi(bx, 0, p.W - 3);
    by = clampi(by, 0, p.H - 3);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float w = wx[a] * wy[b];
            const float dpx = ((float)a - fx) * p.dx;
            const float dpy = ((float)b - fy) * p.dx;
            const float afx = dpx * Q.a00 + dpy * Q.a01;
            const float afy = dpx * Q.a10 + dpy * Q.a11;
            const size_t node = gb + (size_t)(by + b) * (size_t)p.W + (size_t)(bx + a);
            atomic_add_global(&g.mom_x[node], w * (vx * p.mass + afx));
            atomic_add_global(&g.mom_y[node], w * (vy * p.mass + afy));
            atomic_add_global(&g.mass[node], w * p.mass);
        }
    }
}

// G2PKernel + gridToParticle (MPM.cu:178-266): one particle per lane, 9 velocity gathers from HBM.
__global__ __launch_bounds__(kThreads) void k_g2p_reference(DevParams p, Store s, const int32_t* count, Grid g)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count[0]) { return; }
    const int scene = scene_of_key(p, s.key[i]);
    const size_t gb = (size_t)scene * (size_t)p.G;
    float px = s.x[i], py = s.y[i];
    int bx, by;
    float fx, fy, wx[3], wy[3];
    bspline(px, p.inv_dx, (float)p.halo, bx, fx, wx);
    bspline(py, p.inv_dx, (float)p.halo, by, fy, wy);
    bx = clampi(bx, 0, p.W - 3);
    by = clampi(by, 0, p.H - 3);
    float vx = 0.f, vy = 0.f;
    Mat2 B = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float w = wx[a] * wy[b];
            const float dpx = ((float)a - fx) * p.dx;
            const float dpy = ((float)b - fy) * p.dx;
            const float2 gv = g.vel[gb + (size_t)(by + b) * (size_t)p.W + (size_t)(bx + a)];
            vx += w * gv.x;
            vy += w * gv.y;
            const float wvx = w * gv.x, wvy = w * gv.y;
            B.a00 += wvx * dpx;
            B.a01 += wvx * dpy;
            B.a10 += wvy * dpx;
            B.a11 += wvy * dpy;
        }
    }
    const float sB = 4.f * p.inv_dx * p.inv_dx;  // C = 4 inv_dx^2 B, MPM.cu:255
    const Mat2 C = {sB * B.a00, sB * B.a01, sB * B.a10, sB * B.a11};
    Mat2 F = load_F(s, i);
    float J = s.J[i];
    g2p_update(p, s.mat[i], px, py, vx, vy, C, F, J);
    s.x[i] = px;
    s.y[i] = py;
    s.vx[i] = vx;
    s.vy[i] = vy;
    store_C(s, i, C);
    store_F(s, i, F);
    s.J[i] = J;
}

hipError_t launch_p2g_reference(hipStream_t st, const DevParams& p, const Store& s, const int32_t* count, int cap,
                                const Grid& g)
{
    hipLaunchKernelGGL(k_p2g_reference, dim3((unsigned)((cap + kThreads - 1) / kThreads)), dim3(kThreads), 0, st, p,
                       s, count, g);
    return hipGetLastError();
}
hipError_t launch_g2p_reference(hipStream_t st, const DevParams& p, const Store& s, const int32_t* count, int cap,
                                const Grid& g)
{
    hipLaunchKernelGGL(k_g2p_reference, dim3((unsigned)((cap + kThreads - 1) / kThreads)), dim3(kThreads), 0, st, p,
                       s, count, g);
    return hipGetLastError();
}

// ===================================================================================== binning (stable, layered)

// Sort key of every particle from its position: bin key (the scene comes from the previous key) above the
// stencil base node within the bin.
// Also clears the bin ranges of the binning (bins without particles keep start = end = 0), which k_bin_bounds
// fills in later on the same stream.
__global__ __launch_bounds__(kThreads) void k_bin_keys(DevParams p, Store s, const int32_t* count, uint32_t* keys,
                                                     int32_t* bstart, int32_t* bend, int nkeys)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    for (int k = i; k < nkeys; k += gridDim.x * blockDim.x) {
        bstart[k] = 0;
        bend[k] = 0;
    }
    if (i >= count[0]) { return; }
    const uint32_t scene = p.num_scenes > 1 ? s.key[i] / (uint32_t)p.bps : 0u;
    uint32_t node;
    const uint32_t bin = bin_key(p, s.x[i], s.y[i], scene, node);
    keys[i] = (bin << p.node_bits) | node;
}

// Radix upsweep: digit histogram of one tile of kRadixTile keys, written digit-major hist[d * ntiles + t].
// The keys arrive nearly sorted (the previous binning's order), so the 64 keys a wave reads mostly share
// one digit: such a wave adds its count with one LDS atomic instead of 64 on the same counter.
__global__ __launch_bounds__(kThreads) void k_radix_upsweep(const uint32_t* keys, const int32_t* count, int shift,
                                                            int32_t* hist, int ntiles)
{
    __shared__ int32_t lh[kRadixBins];
    const int t = blockIdx.x;
    const int n = count[0];
    for (int d = threadIdx.x; d < kRadixBins; d += kThreads) { lh[d] = 0; }
    __syncthreads();
    const int base = t * kRadixTile;
    if (base < n) {
#pragma unroll 4
        for (int r = 0; r < kRadixItems; ++r) {
            const int idx = base + r * kThreads + threadIdx.x;
            const bool valid = idx < n;
            const int d = valid ? (int)((keys[idx] >> shift) & (kRadixBins - 1)) : 0;
            const int d0 = __shfl(d, 0, 64);
            const uint64_t live = __ballot(valid);
            if (__ballot(valid && d == d0) == live) {
                if ((threadIdx.x & 63) == 0 && live != 0ull) { atomicAdd(&lh[d0], __popcll(live)); }
            } else if (valid) {
                atomicAdd(&lh[d], 1);
            }
        }
    }
    __syncthreads();
    for (int d = threadIdx.x; d < kRadixBins; d += kThreads) { hist[d * ntiles + t] = lh[d]; }
}

// Work-group exclusive scan of kThreads values; returns the prefix of the calling lane, *total the sum.
__device__ __forceinline__ int block_exclusive_scan(int v, int* total)
{
    __shared__ int wsum[kThreads / 64];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) { x += y; }
    }
    if (lane == 63) { wsum[wave] = x; }
    __syncthreads();
    int wpre = 0, tot = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) {
        const int c = wsum[w];
        if (w < wave) { wpre += c; }
        tot += c;
    }
    __syncthreads();
    *total = tot;
    return wpre + x - v;
}

constexpr int kScanItems = 16;
constexpr int kScanSeg = kThreads * kScanItems;

// Scan phase 1: total of each segment of kScanSeg values.
__global__ __launch_bounds__(kThreads) void k_scan_partials(const int32_t* data, int len, int32_t* partials)
{
    const int base = blockIdx.x * kScanSeg + threadIdx.x * kScanItems;
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int i = base + k;
        if (i < len) { s += data[i]; }
    }
    int total;
    block_exclusive_scan(s, &total);
    if (threadIdx.x == 0) { partials[blockIdx.x] = total; }
}
// Scan phase 2: exclusive scan of each segment in place, offset by the sum of the totals of the segments
// before it (each work group adds them up itself: there are few segments, and one launch less).
__global__ __launch_bounds__(kThreads) void k_scan_apply(int32_t* data, int len, const int32_t* partials)
{
    int before = 0;
    for (int b = threadIdx.x; b < (int)blockIdx.x; b += kThreads) { before += partials[b]; }
    int prefix;
    block_exclusive_scan(before, &prefix);
    const int base = blockIdx.x * kScanSeg + threadIdx.x * kScanItems;
    int vals[kScanItems];
    int s = 0;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int i = base + k;
        vals[k] = i < len ? data[i] : 0;
        s += vals[k];
    }
    int total;
    int run = block_exclusive_scan(s, &total) + prefix;
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const int i = base + k;
        if (i < len) { data[i] = run; }
        run += vals[k];
    }
}

static hipError_t exclusive_scan(hipStream_t st, int32_t* data, int len, int32_t* partials)
{
    const int nseg = (len + kScanSeg - 1) / kScanSeg;
    if (nseg == 0) { return hipSuccess; }
    hipLaunchKernelGGL(k_scan_partials, dim3(nseg), dim3(kThreads), 0, st, data, len, partials);
    hipLaunchKernelGGL(k_scan_apply, dim3(nseg), dim3(kThreads), 0, st, data, len, partials);
    return hipGetLastError();
}

// Radix downsweep: stable scatter of one tile.  The tile is consumed in kRadixItems rounds of kThreads
// consecutive keys; inside a round a wave ranks its lanes by digit with 8 ballots (lanes with the same
// digit and a lower lane id come first), the waves of the round are ordered by a per-digit prefix over
// wave counts, and a running per-digit count carries the order across rounds.  Ranks therefore follow
// the input index order exactly: the sort is stable and deterministic.
__global__ __launch_bounds__(kThreads) void k_radix_downsweep(const uint32_t* keys_in, const int32_t* vals_in,
                                                              const int32_t* count, int shift, const int32_t* offsets,
                                                              int ntiles, uint32_t* keys_out, int32_t* vals_out)
{
    constexpr int kWaves = kThreads / 64;
    __shared__ int32_t wcnt[kWaves][kRadixBins];
    __shared__ int32_t wpre[kWaves][kRadixBins];
    __shared__ int32_t run[kRadixBins];
    const int t = blockIdx.x;
    const int n = count[0];
    const int base = t * kRadixTile;
    if (base >= n) { return; }
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int d = threadIdx.x; d < kRadixBins; d += kThreads) {
        run[d] = offsets[d * ntiles + t];
#pragma unroll
        for (int w = 0; w < kWaves; ++w) { wcnt[w][d] = 0; }
    }
    __syncthreads();
    for (int r = 0; r < kRadixItems; ++r) {
        const int idx = base + r * kThreads + threadIdx.x;
        const bool valid = idx < n;
        uint32_t key = 0;
        int32_t val = 0;
        if (valid) {
            key = keys_in[idx];
            val = vals_in ? vals_in[idx] : idx;
        }
        const uint32_t digit = (key >> shift) & (kRadixBins - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < kRadixBits; ++b) {
            const uint64_t m = __ballot((digit >> b) & 1u);
            peers &= ((digit >> b) & 1u) ? m : ~m;
        }
        const int rank = __popcll(peers & lt_mask);
        if (valid && (peers & lt_mask) == 0) { wcnt[wave][digit] = __popcll(peers); }
        __syncthreads();
        for (int d = threadIdx.x; d < kRadixBins; d += kThreads) {
            int acc = run[d];
#pragma unroll
            for (int w = 0; w < kWaves; ++w) {
                wpre[w][d] = acc;
                acc += wcnt[w][d];
                wcnt[w][d] = 0;
            }
            run[d] = acc;
        }
        __syncthreads();
        if (valid) {
            const int pos = wpre[wave][digit] + rank;
            if ((unsigned)pos < (unsigned)n) {  // always true for a consistent histogram
                keys_out[pos] = key;
                vals_out[pos] = val;
            }
        }
    }
}

// Permutation gather of the particle state: dst[j] = src[perm[j]] (keys already in dst.key).  With Full
// false, v and C stay behind: a binning placed after a substep's P2G, whose next reader of v and C is the
// G2P that rewrites them.
// Work groups are dispatched round-robin over the 8 XCDs; the grid is a multiple of 8 and XCD x takes the
// x-th contiguous eighth of the destination, so the source lines of a bin (whose particles stay close to
// their previous positions) are fetched into one XCD's L2 rather than into all eight.
template <bool Full>
__global__ __launch_bounds__(kThreads) void k_gather(Store src, Store dst, const int32_t* perm, const int32_t* count)
{
    const unsigned per = gridDim.x >> 3;
    const unsigned lb = (blockIdx.x & 7u) * per + (blockIdx.x >> 3);
    const int j = (int)(lb * blockDim.x + threadIdx.x);
    const int n = count[0];
    if (j >= n) { return; }
    const int i = perm[j];
    if ((unsigned)i >= (unsigned)n) { return; }
    dst.x[j] = src.x[i];
    dst.y[j] = src.y[i];
    if (Full) {
        dst.vx[j] = src.vx[i];
        dst.vy[j] = src.vy[i];
        dst.c00[j] = src.c00[i];
        dst.c01[j] = src.c01[i];
        dst.c10[j] = src.c10[i];
        dst.c11[j] = src.c11[i];
    }
    dst.f00[j] = src.f00[i];
    dst.f01[j] = src.f01[i];
    dst.f10[j] = src.f10[i];
    dst.f11[j] = src.f11[i];
    dst.J[j] = src.J[i];
    dst.mat[j] = src.mat[i];
    dst.id[j] = src.id[i];
}

// Bin ranges from the sorted keys, bin = key >> shift (bins not present keep start = end = 0).
__global__ __launch_bounds__(kThreads) void k_bin_bounds(const uint32_t* keys, const int32_t* count, int nkeys,
                                                         int shift, int32_t* bstart, int32_t* bend)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int n = count[0];
    if (j >= n) { return; }
    const uint32_t k = keys[j] >> shift;
    if (k >= (uint32_t)nkeys) { return; }
    if (j == 0 || (keys[j - 1] >> shift) != k) { bstart[k] = j; }
    if (j == n - 1 || (keys[j + 1] >> shift) != k) { bend[k] = j + 1; }
}

// Layered order within each bin.  Input: particles sorted stably by bin (the radix passes skip the node
// bits of the key), so inside a bin they keep their index order.  Particle j of node n has rank r = the
// number of particles of node n before it in the bin; it goes to layer min(r, kLayers - 1).  Layers follow
// one another, and inside a layer below the last each node holds one particle, so the 64 lanes of a wave
// mostly start their stencils at distinct nodes and the ranked scatter needs fewer rounds.  Inside a layer
// the nodes follow in checkerboard order: first those with lx + ly even, then those with lx + ly odd, each
// group in node order.  The lanes of a wave then start two cells (or one diagonal step) apart, so particles
// that drift by a cell during the substeps between two sorts still mostly start at distinct nodes.  The
// last layer takes the remaining particles in (checkerboard slot, rank) order.  One work group per bin; the
// result equals the layered order of a stable sort by (bin, node).
// Most layers of the in-bin order: particle of rank r at its node goes to layer min(r, kLayers - 1).  Scenes of
// 16 particles per cell put about 16 at a node, so fewer layers left long runs of one node in the last
// layer (64 scenes of 256^2 cells at 16 per cell: 2.5 scatter rounds per wave with 16 layers, 8 % slower
// ticks).  A bin runs only as many layers as its fullest node needs; the order is the same as with all
// kLayers.
constexpr int kLayers = 24;

// Node (ly * bin_x + lx) at checkerboard slot q of a bin of nn = bin_x * bin_y nodes, bin_x even: slots
// [0, nn/2) are the nodes with lx + ly even, slots [nn/2, nn) those with lx + ly odd, each in node order.
__device__ __forceinline__ int checker_node(int q, int nn, int bin_x)
{
    const int c = q >= (nn >> 1) ? 1 : 0;
    const int r = q - c * (nn >> 1);
    const int hx = bin_x >> 1;
    const int ly = r / hx;
    const int lx = 2 * (r - ly * hx) + ((ly + c) & 1);
    return ly * bin_x + lx;
}
// LDS of k_layer_bins for bins of 2^node_bits nodes (at least 64: the node prefixes run over whole waves);
// 60.3 KB for the largest compiled shape (32 x 16 cells, 512 nodes).
static size_t layer_bins_lds_bytes(int node_bits)
{
    const size_t nn = (size_t)1 << node_bits;
    return sizeof(int) * ((size_t)(kThreads / 64) * nn + nn + (size_t)kLayers * nn + (size_t)kLayers * ((nn + 63) / 64) +
                          (size_t)kLayers + 1);
}
__global__ __launch_bounds__(kThreads) void k_layer_bins(const uint32_t* keys_in, const int32_t* perm_in,
                                                         const int32_t* bstart, const int32_t* bend, int nkeys,
                                                         int node_bits, int bin_x, uint32_t* keys_out,
                                                         int32_t* perm_out)
{
    constexpr int kWaves = kThreads / 64;
    // LDS sized by the launch for nn = 2^node_bits nodes (layer_bins_lds_bytes), so small bins leave room
    // for more work groups per CU:
    //   wrun[w * nn + n]   per wave w and node n: particles before the wave's share, then its running count
    //   cnt[n]             particles of node n
    //   pre[l * nn + n]    layer l < last: nodes before n in the layer; last: particles before n
    //   tot[l * nw + v]    per layer and node wave v (nw = nn / 64): the nodes (or particles) it holds
    //   loff[l]            first position of layer l
    //   nl                 layers this bin uses: min(most particles at one node, kLayers)
    extern __shared__ int lds[];
    const int nn = 1 << node_bits;
    const int nw = (nn + 63) >> 6;
    int* const wrun = lds;
    int* const cnt = wrun + kWaves * nn;
    int* const pre = cnt + nn;
    int* const tot = pre + kLayers * nn;
    int* const loff = tot + kLayers * nw;
    int* const nl_lds = loff + kLayers;
    const int k = blockIdx.x;
    if (k >= nkeys) { return; }
    const int s = bstart[k], e = bend[k];
    if (e <= s) { return; }
    const uint32_t mask = (uint32_t)nn - 1u;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    // each wave ranks a contiguous share of the bin, in whole groups of 64 particles
    const int len = e - s;
    const int gper = (((len + 63) >> 6) + kWaves - 1) / kWaves;
    const int ws = s + min(wave * gper * 64, len);
    const int we = s + min((wave + 1) * gper * 64, len);
    for (int n = threadIdx.x; n < nn; n += kThreads) {
#pragma unroll
        for (int w = 0; w < kWaves; ++w) { wrun[w * nn + n] = 0; }
    }
    if (threadIdx.x == 0) { *nl_lds = 1; }
    __syncthreads();
    for (int j = ws + lane; j < we; j += 64) { atomicAdd(&wrun[wave * nn + (int)(keys_in[j] & mask)], 1); }
    __syncthreads();
    // per node: total, and the exclusive prefix over the waves' shares
    for (int n = threadIdx.x; n < nn; n += kThreads) {
        int c = 0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) {
            const int t = wrun[w * nn + n];
            wrun[w * nn + n] = c;
            c += t;
        }
        cnt[n] = c;
        if (c > 1) { atomicMax(nl_lds, min(c, kLayers)); }
    }
    __syncthreads();
    const int nl = *nl_lds;  // layers [nl, kLayers) would be empty: the order is that of kLayers layers
    // in-wave prefixes over checkerboard slots by ballot (layers below the last) and by a wave scan (last)
    for (int q = threadIdx.x; q < nn; q += kThreads) {
        const int n = checker_node(q, nn, bin_x);
        const int c = cnt[n];
        const int w = q >> 6;
        for (int l = 0; l < nl - 1; ++l) {
            const uint64_t m = __ballot(c > l);
            pre[l * nn + n] = __popcll(m & lt_mask);
            if (lane == 0) { tot[l * nw + w] = __popcll(m); }
        }
        const int over = c > nl - 1 ? c - (nl - 1) : 0;
        int incl = over;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int t = __shfl_up(incl, d, 64);
            if (lane >= d) { incl += t; }
        }
        pre[(nl - 1) * nn + n] = incl - over;
        if (lane == 63) { tot[(nl - 1) * nw + w] = incl; }
    }
    __syncthreads();
    for (int q = threadIdx.x; q < nn; q += kThreads) {
        const int n = checker_node(q, nn, bin_x);
        const int w = q >> 6;
        for (int l = 0; l < nl; ++l) {
            int carry = 0;
            for (int v = 0; v < w; ++v) { carry += tot[l * nw + v]; }
            pre[l * nn + n] += carry;
        }
    }
    if (threadIdx.x == 0) {
        int o = 0;
        for (int l = 0; l < nl; ++l) {
            loff[l] = o;
            for (int v = 0; v < nw; ++v) { o += tot[l * nw + v]; }
        }
    }
    __syncthreads();
    // Each wave walks its share in index order, 64 particles at a time: the rank of a particle is the
    // wave's running count of its node plus the lanes below it with the same node (an exact multisplit,
    // one ballot per node bit); the highest lane of each node then advances the running count.  A wave
    // executes its LDS instructions in order, so the next group reads the advanced counts.
    for (int base = ws; base < we; base += 64) {
        const int j = base + lane;
        const bool valid = j < we;
        const uint32_t key = valid ? keys_in[j] : 0u;
        const int n = (int)(key & mask);
        uint64_t peers = __ballot(valid);
        for (int b = 0; b < node_bits; ++b) {
            const bool bit = (n >> b) & 1;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        int r = 0;
        if (valid) { r = wrun[wave * nn + n] + __popcll(peers & lt_mask); }
        __asm__ __volatile__("" ::: "memory");
        if (valid && (peers >> lane) == 1ull) { wrun[wave * nn + n] = r + 1; }
        __asm__ __volatile__("" ::: "memory");
        const int off = r < nl - 1 ? loff[r] + pre[r * nn + n] : loff[nl - 1] + pre[(nl - 1) * nn + n] + (r - (nl - 1));
        if (valid && off < len) {  // always true for a correct ranking; keeps a wrong one inside the bin
            perm_out[s + off] = perm_in[j];
            keys_out[s + off] = key >> node_bits;
        }
    }
}
// Work items per bin: ceil(count / chunk).
__global__ __launch_bounds__(kThreads) void k_item_counts(const int32_t* bstart, const int32_t* bend, int nkeys,
                                                          int chunk, int32_t* cnt)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nkeys) { return; }
    const int c = bend[k] - bstart[k];
    cnt[k] = (c + chunk - 1) / chunk;
}
// Work item list from the scanned item counts; the last bin also publishes the total.
__global__ __launch_bounds__(kThreads) void k_items_write(const int32_t* bstart, const int32_t* bend,
                                                          const int32_t* off, int nkeys, int chunk, Item* items,
                                                          int32_t* num_items)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nkeys) { return; }
    const int b = bstart[k], e = bend[k];
    const int c = (e - b + chunk - 1) / chunk;
    const int o = off[k];
    for (int q = 0; q < c; ++q) {
        Item it;
        it.key = k;
        it.begin = b + q * chunk;
        it.end = min(b + (q + 1) * chunk, e);
        it.pad = 0;
        items[o + q] = it;
    }
    if (k == nkeys - 1) { num_items[0] = o + c; }
}

hipError_t launch_binning(hipStream_t st, const DevParams& p, const Store& src, const Store& dst,
                          const int32_t* count, const SortScratch& ss, bool full)
{
    const int cap = ss.capacity;
    const unsigned pblocks = (unsigned)((cap + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(k_bin_keys, dim3(pblocks), dim3(kThreads), 0, st, p, src, count, ss.keys_tmp, ss.bin_start,
                       ss.bin_end, ss.num_keys);
    // the passes sort by the bin bits only: k_layer_bins ranks the nodes inside each bin itself
    const int bin_bits = ss.key_bits - p.node_bits;
    const int passes = (bin_bits + kRadixBits - 1) / kRadixBits;
    if (passes <= 0) { return hipErrorInvalidValue; }
    const uint32_t* kin = ss.keys_tmp;
    const int32_t* vin = nullptr;
    for (int pass = 0; pass < passes; ++pass) {
        uint32_t* kout = pass % 2 == 0 ? ss.keys_alt : ss.keys_tmp;
        int32_t* vout = pass % 2 == 0 ? ss.vals_alt : ss.vals_tmp;
        const int shift = p.node_bits + pass * kRadixBits;
        hipLaunchKernelGGL(k_radix_upsweep, dim3(ss.ntiles), dim3(kThreads), 0, st, kin, count, shift, ss.hist,
                           ss.ntiles);
        hipError_t e = exclusive_scan(st, ss.hist, kRadixBins * ss.ntiles, ss.partials);
        if (e != hipSuccess) { return e; }
        hipLaunchKernelGGL(k_radix_downsweep, dim3(ss.ntiles), dim3(kThreads), 0, st, kin, vin, count, shift,
                           ss.hist, ss.ntiles, kout, vout);
        kin = kout;
        vin = vout;
    }
    // kin / vin: keys sorted stably by bin and the permutation; bin ranges, then the layered order
    if (layer_bins_lds_bytes(p.node_bits) > 65536) { return hipErrorInvalidValue; }
    hipLaunchKernelGGL(k_bin_bounds, dim3(pblocks), dim3(kThreads), 0, st, kin, count, ss.num_keys, p.node_bits,
                       ss.bin_start, ss.bin_end);
    hipLaunchKernelGGL(k_layer_bins, dim3((unsigned)ss.num_keys), dim3(kThreads), layer_bins_lds_bytes(p.node_bits), st,
                       kin, vin, ss.bin_start,
                       ss.bin_end, ss.num_keys, p.node_bits, p.bin_x, dst.key, ss.perm);
    if (full) {
        hipLaunchKernelGGL(k_gather<true>, dim3((pblocks + 7u) / 8u * 8u), dim3(kThreads), 0, st, src, dst, ss.perm,
                           count);
    } else {
        hipLaunchKernelGGL(k_gather<false>, dim3((pblocks + 7u) / 8u * 8u), dim3(kThreads), 0, st, src, dst, ss.perm,
                           count);
    }
    const unsigned kblocks = (unsigned)((ss.num_keys + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(k_item_counts, dim3(kblocks), dim3(kThreads), 0, st, ss.bin_start, ss.bin_end, ss.num_keys,
                       ss.chunk, ss.item_cnt);
    hipError_t e = exclusive_scan(st, ss.item_cnt, ss.num_keys, ss.partials);
    if (e != hipSuccess) { return e; }
    hipLaunchKernelGGL(k_items_write, dim3(kblocks), dim3(kThreads), 0, st, ss.bin_start, ss.bin_end, ss.item_cnt,
                       ss.num_keys, ss.chunk, ss.items, ss.num_items);
    return hipGetLastError();
}

#endif  // MPM_MAIN_TU

// ===================================================================================== binned kernels

// Work groups are dispatched round-robin over the 8 XCDs; in one scene, give each XCD a contiguous run of
// work items (neighbouring bins) so the velocity tiles and accumulator flushes of neighbours share that
// XCD's L2.  With several scenes the items go round-robin too: the runs would give whole scenes of unequal
// cost to one XCD (measured per tick with the round-robin order: 64 scenes of 256^2 cells -2.5 %, 16 scenes
// of 2048^2 -2 %, one scene of 1024^2 +2.7 %).
__device__ __forceinline__ int xcd_item(int num_items, bool round_robin, bool& ok)
{
    if (round_robin) {
        ok = (int)blockIdx.x < num_items;
        return (int)blockIdx.x;
    }
    const int per = (num_items + 7) >> 3;
    const int xcd = blockIdx.x & 7;
    const int slot = blockIdx.x >> 3;
    const int it = xcd * per + slot;
    ok = slot < per && it < num_items;
    return it;
}

template <int BX, int BY>
struct TileShape {
    static constexpr int TX = BX + 2 + 2 * kMargin;
    static constexpr int TY = BY + 2 + 2 * kMargin;
    static constexpr int N = TX * TY;
};

__device__ __forceinline__ void item_origin(const DevParams& p, int key, int BX, int BY, int& scene, int& ox, int& oy)
{
    scene = key / p.bps;
    const int r = key - scene * p.bps;
    const int by = r / p.nbx;
    const int bx = r - by * p.nbx;
    ox = bx * BX - kMargin;
    oy = by * BY - kMargin;
}

// Velocity of grid node (gx, gy) of the scene at gb: the grid-update output, or, in the fused substep,
// computed here from the previous substep's accumulators (gridKernel, MPM.cu:131-177, folded into G2P).
__device__ __forceinline__ float2 grid_velocity(const DevParams& p, const Grid& g, size_t gb, int gx, int gy)
{
    const size_t n = gb + (size_t)gy * p.W + gx;
    if (g.src_m) { return node_velocity(p, g.src_x[n], g.src_y[n], g.src_m[n], gx, gy); }
    return g.vel[n];
}

// Two nodes per thread per pass, with all their loads issued before either velocity is formed: a tile of
// N <= 2 kThreads nodes waits for HBM once, not twice.
template <int BX, int BY>
__device__ __forceinline__ void load_vel_tile(const DevParams& p, const Grid& g, size_t gb, int ox, int oy, float2* sv)
{
    using T = TileShape<BX, BY>;
    for (int t0 = threadIdx.x; t0 < T::N; t0 += 2 * kThreads) {
        int gx[2], gy[2];
        bool in[2];
        float mx[2], my[2], m[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int t = t0 + k * kThreads;
            const int ty = t / T::TX;
            const int tx = t - ty * T::TX;
            gx[k] = ox + tx;
            gy[k] = oy + ty;
            in[k] = t < T::N && gx[k] >= 0 && gx[k] < p.W && gy[k] >= 0 && gy[k] < p.H;
            mx[k] = my[k] = m[k] = 0.f;
            if (in[k]) {
                const size_t n = gb + (size_t)gy[k] * p.W + gx[k];
                if (g.src_m) {
                    mx[k] = g.src_x[n];
                    my[k] = g.src_y[n];
                    m[k] = g.src_m[n];
                } else {
                    const float2 v = g.vel[n];
                    mx[k] = v.x;
                    my[k] = v.y;
                }
            }
        }
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int t = t0 + k * kThreads;
            if (t < T::N) {
                float2 v = {0.f, 0.f};
                if (in[k]) { v = g.src_m ? node_velocity(p, mx[k], my[k], m[k], gx[k], gy[k]) : make_float2(mx[k], my[k]); }
                sv[t] = v;
            }
        }
    }
}

// Fused substep: every work group of the launch (also those without a work item) zeroes a slice of the
// accumulators the substep after next scatters into.  Nobody reads or writes them during this launch: the
// previous launch read them, the next one writes the other buffer.
__device__ __forceinline__ void zero_next_accumulators(const Grid& g)
{
    if (!g.zero_x) { return; }
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < g.zero_n; i += stride) {
        g.zero_x[i] = 0.f;
        g.zero_y[i] = 0.f;
        g.zero_m[i] = 0.f;
    }
}

// LDS accumulators of P2G: per wave one tile of N nodes holding momentum x, momentum y and the weight sum
// (mass / particle mass: every particle has the same mass, applied once per node at the flush).  Each
// wave owns its tile, so no two waves ever update the same LDS word; the tiles are summed in a fixed wave
// order at the flush.  Layout (PlaneTiles): momentum in a float2 plane and the weight sum in a float plane,
// so a node update is one b64 and one b32 read-add-write.  Measured on gfx950 for a 9-node stencil of 64
// lanes on distinct start nodes (4 waves per SIMD): 163 CU cycles per wave, where float4 nodes updated
// with b96 accesses take 244 (start nodes in row order) to 295 (checkerboard order) and with b128 accesses
// 176 to 225.  A byte plane holds the start-node claims of the scatter's ranking.
template <int N>
struct PlaneTiles {
    float2 mom[N * kAccTiles];
    float mass[N * kAccTiles];
    uint8_t owner[N * kAccTiles];  // scratch of the ranking in scatter (written before every read)
    // Nodes relative to a start node: the stencil offsets become immediate offsets of the LDS instructions.
    struct Cursor {
        float2* mom;
        float* mass;
        // both planes are read before either is written (one LDS round trip per node, not two)
        __device__ __forceinline__ void add(int k, float x, float y, float m) const
        {
            float2 v = mom[k];
            float w = mass[k];
            v.x += x;
            v.y += y;
            w += m;
            mom[k] = v;
            mass[k] = w;
        }
    };
    struct Wave {
        float2* mom;
        float* mass;
        uint8_t* owner;
        __device__ __forceinline__ Cursor at(int node) const { return {mom + node, mass + node}; }
    };
    __device__ __forceinline__ Wave wave(int w) { return {mom + w * N, mass + w * N, owner + w * N}; }
    __device__ __forceinline__ void zero(int t)
    {
        mom[t] = make_float2(0.f, 0.f);
        mass[t] = 0.f;
    }
    // (momentum x, momentum y, mass) of node t summed over the waves in wave order
    __device__ __forceinline__ float3 sum(int t) const
    {
        float3 v = make_float3(mom[t].x, mom[t].y, mass[t]);
#pragma unroll
        for (int w = 1; w < kAccTiles; ++w) {
            v.x += mom[w * N + t].x;
            v.y += mom[w * N + t].y;
            v.z += mass[w * N + t];
        }
        return v;
    }
};
// Minimum waves per SIMD requested from the register allocator for the binned kernels of one tile shape:
// the target (7 for the plain fused kernel, 5 for the learned kernels), capped by what the work group's LDS
// allows on a CU (160 KB; one work group of 4 waves holds one wave per SIMD).
template <int BX, int BY, bool kLearned, int kTarget>
struct WavesPerSimd {
    static constexpr int N = TileShape<BX, BY>::N;
    static constexpr int lds = (int)(sizeof(PlaneTiles<N>) + sizeof(float2) * N) +
                               (kLearned ? (int)sizeof(ClosurePacked) : 0);
    static constexpr int by_lds = (160 * 1024) / lds;
    static constexpr int value = by_lds < kTarget ? (by_lds < 1 ? 1 : by_lds) : kTarget;
};

template <int BX, int BY, class Tiles>
__device__ __forceinline__ void zero_acc_tiles(Tiles& acc)
{
    using T = TileShape<BX, BY>;
    for (int t = threadIdx.x; t < T::N * kAccTiles; t += kThreads) { acc.zero(t); }
}

// One HBM atomic per touched tile node and component (the LDS tiles already reduced every particle of the
// work item).
template <int BX, int BY, class Tiles>
__device__ __forceinline__ void flush_acc_tiles(const DevParams& p, const Grid& g, size_t gb, int ox, int oy,
                                                const Tiles& acc)
{
    using T = TileShape<BX, BY>;
    for (int t = threadIdx.x; t < T::N; t += kThreads) {
        const float3 v = acc.sum(t);
        if (v.z == 0.f && v.x == 0.f && v.y == 0.f) { continue; }
        const int ty = t / T::TX;
        const int tx = t - ty * T::TX;
        const int gx = ox + tx, gy = oy + ty;
        if (gx < 0 || gx >= p.W || gy < 0 || gy >= p.H) { continue; }
        const size_t node = gb + (size_t)gy * p.W + gx;
        atomic_add_global(&g.mom_x[node], v.x);
        atomic_add_global(&g.mom_y[node], v.y);
        atomic_add_global(&g.mass[node], v.z * p.mass);
    }
}

// Gather of one particle (MPM.cu:227-251) from the LDS velocity tile, or from HBM when it drifted out of
// the tile.  v = sum w_ab v_ab, and with d_ab = ((a - fx) dx, (b - fy) dx) the affine moment
// B = sum w_ab v_ab (x) d_ab = dx (sum w_ab v_ab (a, b) - v (x) f): two first moments replace the
// per-node outer products (each column of three nodes is reduced with wy first).  The gather returns
// C = 4 inv_dx^2 B (MPM.cu:255) = 4 inv_dx (sum w_ab v_ab (a, b) - v (x) f), one scale per entry.
template <class Load>
__device__ __forceinline__ void gather_moments(const float wx[3], const float wy[3], Load load, float& sx, float& sy,
                                               float& ax, float& ay, float& bx, float& by)
{
    sx = sy = ax = ay = bx = by = 0.f;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float cx = 0.f, cy = 0.f, dx = 0.f, dy = 0.f;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float2 gv = load(a, b);
            cx = fmaf(wy[b], gv.x, cx);
            cy = fmaf(wy[b], gv.y, cy);
            if (b > 0) {
                const float wb = wy[b] * (float)b;
                dx = fmaf(wb, gv.x, dx);
                dy = fmaf(wb, gv.y, dy);
            }
        }
        sx = fmaf(wx[a], cx, sx);
        sy = fmaf(wx[a], cy, sy);
        bx = fmaf(wx[a], dx, bx);
        by = fmaf(wx[a], dy, by);
        if (a > 0) {
            const float wa = wx[a] * (float)a;
            ax = fmaf(wa, cx, ax);
            ay = fmaf(wa, cy, ay);
        }
    }
}

// Gather from HBM for a particle that drifted out of its tile (rare): rolled loops keep its registers
// from counting against the tiled path.  Same sums as gather_moments, same order.
__device__ __forceinline__ void gather_moments_hbm(const DevParams& p, const Grid& g, size_t gb, int bx, int by,
                                                   const float wx[3], const float wy[3], float& sx, float& sy,
                                                   float& ax, float& ay, float& cx, float& cy)
{
    sx = sy = ax = ay = cx = cy = 0.f;
#pragma unroll 1
    for (int a = 0; a < 3; ++a) {
        float ux = 0.f, uy = 0.f, dx = 0.f, dy = 0.f;
#pragma unroll 1
        for (int b = 0; b < 3; ++b) {
            const float2 gv = grid_velocity(p, g, gb, bx + a, by + b);
            const float w = b == 0 ? wy[0] : (b == 1 ? wy[1] : wy[2]);
            ux = fmaf(w, gv.x, ux);
            uy = fmaf(w, gv.y, uy);
            if (b > 0) {
                const float wb = w * (float)b;
                dx = fmaf(wb, gv.x, dx);
                dy = fmaf(wb, gv.y, dy);
            }
        }
        const float wxa = a == 0 ? wx[0] : (a == 1 ? wx[1] : wx[2]);
        sx = fmaf(wxa, ux, sx);
        sy = fmaf(wxa, uy, sy);
        cx = fmaf(wxa, dx, cx);
        cy = fmaf(wxa, dy, cy);
        if (a > 0) {
            const float wa = wxa * (float)a;
            ax = fmaf(wa, ux, ax);
            ay = fmaf(wa, uy, ay);
        }
    }
}

template <int BX, int BY>
__device__ __forceinline__ void gather(const DevParams& p, const Grid& g, size_t gb, int ox, int oy, const float2* sv,
                                       float px, float py, float& vx, float& vy, Mat2& C)
{
    using T = TileShape<BX, BY>;
    int bx, by;
    float fx, fy, wx[3], wy[3];
    bspline(px, p.inv_dx, (float)p.halo, bx, fx, wx);
    bspline(py, p.inv_dx, (float)p.halo, by, fy, wy);
    bx = clampi(bx, 0, p.W - 3);
    by = clampi(by, 0, p.H - 3);
    const int lx = bx - ox, ly = by - oy;
    const bool inside = (unsigned)lx <= (unsigned)(T::TX - 3) && (unsigned)ly <= (unsigned)(T::TY - 3);
    float sx, sy, mx, my, nx, ny;
    if (inside) {
        const float2* t0 = sv + ly * T::TX + lx;
        gather_moments(wx, wy, [&](int a, int b) { return t0[b * T::TX + a]; }, sx, sy, mx, my, nx, ny);
    } else {
        gather_moments_hbm(p, g, gb, bx, by, wx, wy, sx, sy, mx, my, nx, ny);
    }
    vx = sx;
    vy = sy;
    // C = 4 inv_dx^2 B with B = dx (first moments - v (x) f): one scale 4 inv_dx per entry
    const float sC = 4.f * p.inv_dx;
    C = {sC * (mx - fx * sx), sC * (nx - fy * sx), sC * (my - fx * sy), sC * (ny - fy * sy)};
}

// Scatter of one particle into its wave's LDS accumulator tile, without atomics, in rounds in which the
// updating lanes start their stencils at distinct nodes, so for each of the 9 offsets their targets are
// distinct and a plain read-add-write is race free (the compiler barriers keep every read after the
// previous write of the same wave; LDS executes a wave's instructions in order).  Lanes are ranked among
// those starting at the same node (rank = lanes below with the same start node, by an exact multisplit,
// one ballot per bit, over a 6-bit name of the node that a claim in an LDS byte plane provides; skipped
// when the claim shows every start node distinct); round r takes rank r, so the summation order is the
// particle order.  A particle that drifted out of the tile scatters to HBM atomically.
template <int BX, int BY, class WaveTile>
__device__ __forceinline__ void scatter(const DevParams& p, const Grid& g, size_t gb, int ox, int oy,
                                        const WaveTile& wtile, float px, float py, float vx, float vy, const Mat2& Q,
                                        bool active = true)
{
    using T = TileShape<BX, BY>;
    int bx, by;
    float fx, fy, wx[3], wy[3];
    bspline(px, p.inv_dx, (float)p.halo, bx, fx, wx);
    bspline(py, p.inv_dx, (float)p.halo, by, fy, wy);
    bx = clampi(bx, 0, p.W - 3);
    by = clampi(by, 0, p.H - 3);
    const int lx = bx - ox, ly = by - oy;
    const bool inside = active && (unsigned)lx <= (unsigned)(T::TX - 3) && (unsigned)ly <= (unsigned)(T::TY - 3);
    // contribution to stencil node (a, b): w_ab (m v + Q d_ab, m) with d_ab = ((a - fx) dx, (b - fy) dx),
    // i.e. w_ab (A0 + a dx Q.col0 + b dx Q.col1, m), A0 = m v - dx Q f, built incrementally over a and b.
    const float qax = p.dx * Q.a00, qay = p.dx * Q.a10, qbx = p.dx * Q.a01, qby = p.dx * Q.a11;
    const float a0x = vx * p.mass - (qax * fx + qbx * fy);
    const float a0y = vy * p.mass - (qay * fx + qby * fy);
    // The mass contribution is m w_ab with one mass for every particle: the tiles accumulate w_ab and the
    // flush multiplies by m once per node.
    float cx[9], cy[9], cm[9];
    float tax = a0x, tay = a0y;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float tx = tax, ty = tay;
#pragma unroll
        for (int b = 0; b < 3; ++b) {
            const float w = wx[a] * wy[b];
            cx[a * 3 + b] = w * tx;
            cy[a * 3 + b] = w * ty;
            cm[a * 3 + b] = w;
            tx += qbx;
            ty += qby;
        }
        tax += qax;
        tay += qay;
    }
    if (active && !inside) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                const size_t node = gb + (size_t)(by + b) * p.W + (bx + a);
                atomic_add_global(&g.mom_x[node], cx[a * 3 + b]);
                atomic_add_global(&g.mom_y[node], cy[a * 3 + b]);
                atomic_add_global(&g.mass[node], cm[a * 3 + b] * p.mass);
            }
        }
    }
    const int base = ly * T::TX + lx;
    const int lane = threadIdx.x & 63;
    const auto cur = wtile.at(base);
    // Every lane writes its index into the owner byte of its start node and reads the byte back: lanes
    // with the same start node read the same index (that of the lane whose write landed last), lanes
    // with different start nodes different ones.  When no lane reads another's index all start nodes are
    // distinct and every rank is 0; otherwise the index read back is an exact 6-bit name of the start
    // node, and the multisplit runs over it.
    uint8_t* own = wtile.owner + base;
    if (inside) { *own = (uint8_t)lane; }
    __asm__ __volatile__("" ::: "memory");
    const int name = inside ? (int)*own : lane;
    __asm__ __volatile__("" ::: "memory");
    int rank = 0;
    const bool conflict = __ballot(name != lane) != 0ull;  // wave-uniform
    if (conflict) {
        // peers: inside lanes whose name equals this lane's, bit by bit: with m = the lanes whose bit k
        // is set and B = 0 when this lane's bit k is set, all ones otherwise, m ^ B selects m or ~m
        // (the form takes one SGPR operand per VALU instruction)
        const uint64_t in = __ballot(inside);
        uint32_t plo = (uint32_t)in, phi = (uint32_t)(in >> 32);
        const int nn = ~name;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
            const int B = __builtin_amdgcn_sbfe(nn, k, 1);  // bit k of ~name, sign-extended: 0 or -1
            const uint64_t m = __ballot(B == 0);
            plo &= (uint32_t)m ^ (uint32_t)B;
            phi &= (uint32_t)(m >> 32) ^ (uint32_t)B;
        }
        rank = (int)__builtin_amdgcn_mbcnt_hi(phi, __builtin_amdgcn_mbcnt_lo(plo, 0u));  // peers below
    }
    for (int r = 0;; ++r) {
        if (inside && rank == r) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
#pragma unroll
                for (int b = 0; b < 3; ++b) {
                    cur.add(b * T::TX + a, cx[a * 3 + b], cy[a * 3 + b], cm[a * 3 + b]);
                    __asm__ __volatile__("" ::: "memory");
                }
            }
        }
        if (!conflict || __ballot(inside && rank > r) == 0ull) { break; }
    }
}

// P2G over binned particles: one work item (<= chunk particles of one bin) per work group.
template <int BX, int BY>
__global__ __launch_bounds__(kThreads) void k_p2g_tiled(DevParams p, Store s, Grid g, const Item* items,
                                                        const int32_t* num_items, const int32_t* count)
{
    using T = TileShape<BX, BY>;
    __shared__ PlaneTiles<T::N> acc;
    zero_next_accumulators(g);
    bool ok;
    const int it = xcd_item(num_items[0], p.num_scenes > 1, ok);
    if (!ok) { return; }
    Item item = items[it];
    item.end = min(item.end, count[0]);
    int scene, ox, oy;
    item_origin(p, item.key, BX, BY, scene, ox, oy);
    const size_t gb = (size_t)scene * (size_t)p.G;
    zero_acc_tiles<BX, BY>(acc);
    __syncthreads();
    const auto wtile = acc.wave(__builtin_amdgcn_readfirstlane(threadIdx.x >> 6));  // wave-uniform
    for (int j = item.begin + threadIdx.x; j < item.end; j += kThreads) {
        const uint32_t o = (uint32_t)j * 4u;
        const Mat2 Q = affine_q(p, ldo(s.mat, (uint32_t)j), ld_C(s, o), ld_F(s, o), ldo(s.J, o));
        scatter<BX, BY>(p, g, gb, ox, oy, wtile, ldo(s.x, o), ldo(s.y, o), ldo(s.vx, o), ldo(s.vy, o), Q);
    }
    __syncthreads();
    flush_acc_tiles<BX, BY>(p, g, gb, ox, oy, acc);
}

// G2P over binned particles, velocities staged in an LDS tile.
template <int BX, int BY>
__global__ __launch_bounds__(kThreads) void k_g2p_tiled(DevParams p, Store s, Grid g, const Item* items,
                                                        const int32_t* num_items, const int32_t* count)
{
    using T = TileShape<BX, BY>;
    __shared__ float2 sv[T::N];
    bool ok;
    const int it = xcd_item(num_items[0], p.num_scenes > 1, ok);
    if (!ok) { return; }
    Item item = items[it];
    item.end = min(item.end, count[0]);
    int scene, ox, oy;
    item_origin(p, item.key, BX, BY, scene, ox, oy);
    const size_t gb = (size_t)scene * (size_t)p.G;
    load_vel_tile<BX, BY>(p, g, gb, ox, oy, sv);
    __syncthreads();
    for (int j = item.begin + threadIdx.x; j < item.end; j += kThreads) {
        const uint32_t o = (uint32_t)j * 4u;
        float px = ldo(s.x, o), py = ldo(s.y, o), vx, vy;
        const int mat = ldo(s.mat, (uint32_t)j);
        Mat2 C, F = ld_F(s, o);
        float J = ldo(s.J, o);
        gather<BX, BY>(p, g, gb, ox, oy, sv, px, py, vx, vy, C);
        g2p_update(p, mat, px, py, vx, vy, C, F, J);
        sto(s.x, o, px);
        sto(s.y, o, py);
        sto(s.vx, o, vx);
        sto(s.vy, o, vy);
        st_C(s, o, C);
        st_F(s, o, F);
        sto(s.J, o, J);
    }
}

// G2P of substep s fused with P2G of substep s+1: the particle is read once (x, F, J, material), gathers
// from the LDS velocity tile of grid s, is updated, and scatters into the LDS accumulator tiles of grid
// s+1, flushed with one HBM atomic per touched node.  v and C stay in registers (they are recomputed
// by every G2P); they are written only by the last G2P of a segment (k_g2p_tiled).
template <int BX, int BY>
// At least 7 waves per SIMD (at most 72 VGPRs): more resident waves hide the latency of the scatter's
// dependent LDS read-add-write chain (8 would spill to scratch).
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WavesPerSimd<BX, BY, false, 7>::value))) void
k_g2p2g(
    DevParams p, Store s, Grid g, const Item* items, const int32_t* num_items, const int32_t* count)
{
    using T = TileShape<BX, BY>;
    __shared__ float2 sv[T::N];
    __shared__ PlaneTiles<T::N> acc;
    zero_next_accumulators(g);
    bool ok;
    const int it = xcd_item(num_items[0], p.num_scenes > 1, ok);
    if (!ok) { return; }
    Item item = items[it];
    item.end = min(item.end, count[0]);
    int scene, ox, oy;
    item_origin(p, item.key, BX, BY, scene, ox, oy);
    const size_t gb = (size_t)scene * (size_t)p.G;
    load_vel_tile<BX, BY>(p, g, gb, ox, oy, sv);
    zero_acc_tiles<BX, BY>(acc);
    __syncthreads();
    const auto wtile = acc.wave(__builtin_amdgcn_readfirstlane(threadIdx.x >> 6));  // wave-uniform
    // The state of the lane's next particle is loaded one iteration ahead, so an iteration does not wait for
    // HBM: all of it when the register budget of the tile shape allows (at most 6 waves per SIMD, 84+ VGPRs),
    // else the position only (the gather needs it first).
    constexpr bool kAll = WavesPerSimd<BX, BY, false, 7>::value <= 6;
    int j = item.begin + threadIdx.x;
    float nx = 0.f, ny = 0.f, nJ = 1.f;
    int nmat = 0;
    Mat2 nF = {1.f, 0.f, 0.f, 1.f};
    if (j < item.end) {
        nx = ldo(s.x, (uint32_t)j * 4u);
        ny = ldo(s.y, (uint32_t)j * 4u);
        if (kAll) {
            nmat = ldo(s.mat, (uint32_t)j);
            nF = ld_F(s, (uint32_t)j * 4u);
            nJ = ldo(s.J, (uint32_t)j * 4u);
        }
    }
    for (; j < item.end; j += kThreads) {
        const uint32_t o = (uint32_t)j * 4u;
        float px = nx, py = ny, vx, vy;
        int mat = nmat;
        Mat2 C, F = nF;
        float J = nJ;
        if (!kAll) {
            mat = ldo(s.mat, (uint32_t)j);
            F = ld_F(s, o);
            J = ldo(s.J, o);
        }
        if (j + kThreads < item.end) {
            nx = ldo(s.x, o + 4u * kThreads);
            ny = ldo(s.y, o + 4u * kThreads);
            if (kAll) {
                nmat = ldo(s.mat, (uint32_t)(j + kThreads));
                nF = ld_F(s, o + 4u * kThreads);
                nJ = ldo(s.J, o + 4u * kThreads);
            }
        }
        gather<BX, BY>(p, g, gb, ox, oy, sv, px, py, vx, vy, C);
        SnowHint snow = {1.f, 0.f, 1.f};  // a snow particle's rotation and hardening, from its return mapping
        g2p_update(p, mat, px, py, vx, vy, C, F, J, &snow);
        sto(s.x, o, px);
        sto(s.y, o, py);
        st_F(s, o, F);
        sto(s.J, o, J);
        const Mat2 Q = affine_q(p, mat, C, F, J, &snow);
        scatter<BX, BY>(p, g, gb, ox, oy, wtile, px, py, vx, vy, Q);
    }
    __syncthreads();
    flush_acc_tiles<BX, BY>(p, g, gb, ox, oy, acc);
}

// Affine momentum with the learned closure: jelly and snow particles inside the trained range take
// tau = mu * MLP(features); fluid and out-of-range particles the analytic stress.  Wave-uniform: the
// network runs once per wave when any lane needs it (all 64 lanes must be active).
__device__ __forceinline__ Mat2 affine_q_learned(const DevParams& p, const ClosurePacked* cl, int mat, const Mat2& C,
                                                 const Mat2& F, float J)
{
    float x[4] = {0.f, 0.f, 0.f, 0.f};
    bool use = false;
    if (mat != MPM_MAT_FLUID) { use = closure_features(F, mat == MPM_MAT_SNOW ? J : 1.f, x); }
    if (!use) { x[0] = x[1] = x[2] = x[3] = 0.f; }  // finite inputs in every lane: the network's table index
    float y[3] = {0.f, 0.f, 0.f};
    if (__ballot(use) != 0ull) { closure_eval_wave(cl, x, y); }
    if (use) {
        const float D = p.dt * p.vol * 4.f * p.inv_dx * p.inv_dx;
        const float t00 = p.mu * y[0], t01 = p.mu * y[1], t11 = p.mu * y[2];
        return {p.mass * C.a00 - D * t00, p.mass * C.a01 - D * t01, p.mass * C.a10 - D * t01,
                p.mass * C.a11 - D * t11};
    }
    return affine_q(p, mat, C, F, J);
}

// k_p2g_tiled with the learned closure.  Every wave runs the network with all 64 lanes: a wave's trip count
// is uniform over its lanes (lanes past the item's end ride along invalid), and a wave stops when its
// first lane passes the end, so no wave evaluates a network for particles it does not have.
template <int BX, int BY>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WavesPerSimd<BX, BY, true, 5>::value))) void
k_p2g_tiled_learned(DevParams p, Store s, Grid g, const Item* items,
                                                                const int32_t* num_items, const int32_t* count)
{
    using T = TileShape<BX, BY>;
    __shared__ PlaneTiles<T::N> acc;
    __shared__ ClosurePacked cl;
    zero_next_accumulators(g);
    bool ok;
    const int it = xcd_item(num_items[0], p.num_scenes > 1, ok);
    if (!ok) { return; }
    Item item = items[it];
    item.end = min(item.end, count[0]);
    if (item.end <= item.begin) { return; }
    int scene, ox, oy;
    item_origin(p, item.key, BX, BY, scene, ox, oy);
    const size_t gb = (size_t)scene * (size_t)p.G;
    zero_acc_tiles<BX, BY>(acc);
    closure_stage(p.cl, &cl);
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const auto wtile = acc.wave(wave);
    const int lane = threadIdx.x & 63;
    for (int base = item.begin + 64 * wave; base < item.end; base += kThreads) {
        const int j = base + lane;
        const bool valid = j < item.end;
        const uint32_t o = (uint32_t)(valid ? j : item.end - 1) * 4u;
        const int mat = ldo(s.mat, o >> 2);
        const Mat2 Q = affine_q_learned(p, &cl, mat, ld_C(s, o), ld_F(s, o), ldo(s.J, o));
        scatter<BX, BY>(p, g, gb, ox, oy, wtile, ldo(s.x, o), ldo(s.y, o), ldo(s.vx, o), ldo(s.vy, o), Q, valid);
    }
    __syncthreads();
    flush_acc_tiles<BX, BY>(p, g, gb, ox, oy, acc);
}

// k_g2p2g with the learned closure: the P2G stress of every wave comes from one BF16 MFMA evaluation of
// the network for its 64 particles (per-wave trip counts as in k_p2g_tiled_learned).
template <int BX, int BY>
__global__ __launch_bounds__(kThreads) __attribute__((amdgpu_waves_per_eu(WavesPerSimd<BX, BY, true, 5>::value))) void
k_g2p2g_learned(DevParams p, Store s, Grid g, const Item* items,
                                                            const int32_t* num_items, const int32_t* count)
{
    using T = TileShape<BX, BY>;
    __shared__ float2 sv[T::N];
    __shared__ PlaneTiles<T::N> acc;
    __shared__ ClosurePacked cl;
    zero_next_accumulators(g);
    bool ok;
    const int it = xcd_item(num_items[0], p.num_scenes > 1, ok);
    if (!ok) { return; }
    Item item = items[it];
    item.end = min(item.end, count[0]);
    if (item.end <= item.begin) { return; }
    int scene, ox, oy;
    item_origin(p, item.key, BX, BY, scene, ox, oy);
    const size_t gb = (size_t)scene * (size_t)p.G;
    load_vel_tile<BX, BY>(p, g, gb, ox, oy, sv);
    zero_acc_tiles<BX, BY>(acc);
    closure_stage(p.cl, &cl);
    __syncthreads();
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
    const auto wtile = acc.wave(wave);
    const int lane = threadIdx.x & 63;
    // With 5 waves per SIMD (16x8 and 8x8 tiles) the position of the lane's next particle is loaded one
    // iteration ahead (the gather needs it first; lanes past the item's end take its last particle, as in
    // the iteration itself): 168.0 -> 166.7 us per launch on the 4,194,304-particle jelly + snow scene.  The
    // 16x16 shape (3 waves per SIMD, 130 VGPRs) keeps the plain load: there the prefetch led the compiler to
    // a 100-VGPR schedule that serialises the SiLU table reads, 8 % slower.
    constexpr bool kPrefetch = WavesPerSimd<BX, BY, true, 5>::value >= 5;
    float nx = 0.f, ny = 0.f;
    if (kPrefetch) {
        const uint32_t first = (uint32_t)min(item.begin + 64 * wave + lane, item.end - 1) * 4u;
        nx = ldo(s.x, first);
        ny = ldo(s.y, first);
    }
    for (int base = item.begin + 64 * wave; base < item.end; base += kThreads) {
        const int j = base + lane;
        const bool valid = j < item.end;
        const uint32_t o = (uint32_t)(valid ? j : item.end - 1) * 4u;
        float px = nx, py = ny, vx, vy;
        if (!kPrefetch) {
            px = ldo(s.x, o);
            py = ldo(s.y, o);
        } else if (base + kThreads < item.end) {
            const uint32_t on = (uint32_t)min(j + kThreads, item.end - 1) * 4u;
            nx = ldo(s.x, on);
            ny = ldo(s.y, on);
        }
        const int mat = ldo(s.mat, o >> 2);
        Mat2 C, F = ld_F(s, o);
        float J = ldo(s.J, o);
        gather<BX, BY>(p, g, gb, ox, oy, sv, px, py, vx, vy, C);
        g2p_update(p, mat, px, py, vx, vy, C, F, J);
        if (valid) {
            sto(s.x, o, px);
            sto(s.y, o, py);
            st_F(s, o, F);
            sto(s.J, o, J);
        }
        const Mat2 Q = affine_q_learned(p, &cl, mat, C, F, J);
        scatter<BX, BY>(p, g, gb, ox, oy, wtile, px, py, vx, vy, Q, valid);
    }
    __syncthreads();
    flush_acc_tiles<BX, BY>(p, g, gb, ox, oy, acc);
}

#ifdef MPM_MAIN_TU
int shape_index(int bx, int by)
{
    for (int k = 0; k < kNumShapes; ++k) {
        if (kShapeX[k] == bx && kShapeY[k] == by) { return k; }
    }
    return -1;
}

#endif  // MPM_MAIN_TU

static inline unsigned item_blocks(const SortScratch& ss) { return (unsigned)(((ss.max_items + 7) / 8) * 8); }

#define MPM_DISPATCH_SHAPE(KERNEL)                                                                              \
    switch (shape) {                                                                                            \
    case 0: hipLaunchKernelGGL((KERNEL<8, 8>), dim3(nb), dim3(kThreads), 0, st, p, s, g, ss.items, ss.num_items, count); \
        break;                                                                                                  \
    case 1: hipLaunchKernelGGL((KERNEL<16, 8>), dim3(nb), dim3(kThreads), 0, st, p, s, g, ss.items, ss.num_items, count); \
        break;                                                                                                  \
    case 2: hipLaunchKernelGGL((KERNEL<16, 16>), dim3(nb), dim3(kThreads), 0, st, p, s, g, ss.items, ss.num_items, count); \
        break;                                                                                                  \
    case 3: hipLaunchKernelGGL((KERNEL<32, 16>), dim3(nb), dim3(kThreads), 0, st, p, s, g, ss.items, ss.num_items, count); \
        break;                                                                                                  \
    default: return hipErrorInvalidValue;                                                                       \
    }

#ifdef MPM_MAIN_TU
hipError_t launch_p2g_tiled(hipStream_t st, int shape, const DevParams& p, const Store& s, const Grid& g,
                            const SortScratch& ss, const int32_t* count)
{
    const unsigned nb = item_blocks(ss);
    if (p.closure == MPM_CLOSURE_LEARNED) {
        MPM_DISPATCH_SHAPE(k_p2g_tiled_learned)
        return hipGetLastError();
    }
    MPM_DISPATCH_SHAPE(k_p2g_tiled)
    return hipGetLastError();
}
hipError_t launch_g2p_tiled(hipStream_t st, int shape, const DevParams& p, const Store& s, const Grid& g,
                            const SortScratch& ss, const int32_t* count)
{
    const unsigned nb = item_blocks(ss);
    MPM_DISPATCH_SHAPE(k_g2p_tiled)
    return hipGetLastError();
}
hipError_t launch_g2p2g(hipStream_t st, int shape, const DevParams& p, const Store& s, const Grid& g,
                        const SortScratch& ss, const int32_t* count)
{
    const unsigned nb = item_blocks(ss);
    if (p.closure == MPM_CLOSURE_LEARNED) {
        MPM_DISPATCH_SHAPE(k_g2p2g_learned)
        return hipGetLastError();
    }
    return launch_g2p2g_plain(st, shape, p, s, g, ss, count);
}
#else
hipError_t launch_g2p2g_plain(hipStream_t st, int shape, const DevParams& p, const Store& s, const Grid& g,
                              const SortScratch& ss, const int32_t* count)
{
    const unsigned nb = item_blocks(ss);
    MPM_DISPATCH_SHAPE(k_g2p2g)
    return hipGetLastError();
}
#endif  // MPM_MAIN_TU

#ifdef MPM_MAIN_TU
// ===================================================================================== output

__device__ OutView g_out_view;

// MPMOutputKernel (MPM.cu:302-322): positions and velocities into the current output slot, in the
// caller's particle order, with per-chunk particle counts; count[1] = count[0] afterwards.
__global__ __launch_bounds__(kThreads) void k_output(Store s, int32_t* count)
{
    const int n = count[0];
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) { return; }
    const OutView ov = g_out_view;
    const int i = s.id[j];
    if ((unsigned)i >= (unsigned)n) { return; }
    reinterpret_cast<float2*>(ov.pos)[i] = make_float2(s.x[j], s.y[j]);
    reinterpret_cast<float2*>(ov.vel)[i] = make_float2(s.vx[j], s.vy[j]);
    const int oc = ov.out_chunk;
    const int chunk = i / oc;
    if (i + 1 == n) {
        ov.num_atoms[chunk] = (i % oc) + 1;
        count[1] = n;
    } else if ((i % oc) + 1 == oc) {
        ov.num_atoms[chunk] = oc;
    }
}

hipError_t launch_output(hipStream_t st, const Store& s, int32_t* count, int cap)
{
    hipLaunchKernelGGL(k_output, dim3((unsigned)((cap + kThreads - 1) / kThreads)), dim3(kThreads), 0, st, s, count);
    return hipGetLastError();
}

hipError_t set_output_view_async(hipStream_t st, const OutView* host_pinned)
{
    return hipMemcpyToSymbolAsync(HIP_SYMBOL(g_out_view), host_pinned, sizeof(OutView), 0, hipMemcpyHostToDevice, st);
}

// ===================================================================================== emitter

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw 2011, "Parallel random numbers: as easy as 1, 2, 3").
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint2 k)
{
    constexpr uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
        const uint32_t hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
        c = make_uint4(hi1 ^ c.y ^ k.x, lo1, hi0 ^ c.w ^ k.y, lo0);
        k.x += W0;
        k.y += W1;
    }
    return c;
}

// MPMEmitterKernel (MPM.cu:380-438): a width x rows lattice trailing the nozzle in the stream frame,
// Gaussian position jitter (Box-Muller on two Philox outputs), v = velocity / tick, C = 0, F = I, J = 1.
// Appended after count[1]; a wave64 ballot counts the lanes that fit the capacity into count[0].
__global__ __launch_bounds__(kThreads) void k_emitter(DevParams p, Store s, int32_t* count, int cap, float ox,
                                                      float oy, float velx, float vely, int width, float spacing,
                                                      float jitter, int n_emit, uint64_t seed, float inv_tick)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i < n_emit;
    const float v_sqr = velx * velx + vely * vely;
    const float inv_v = (v_sqr > 1e-12f) ? rsqrtf(v_sqr) : 0.f;
    const float tx = velx * inv_v, ty = vely * inv_v;
    const float nx = -ty, ny = tx;
    float px = 0.f, py = 0.f;
    int idx = 0;
    if (active) {
        const int col = i % width;
        const int row = i / width;
        const float across = ((float)col - 0.5f * (float)(width - 1)) * spacing;
        const float along = -((float)row + 0.5f) * spacing;
        px = ox + nx * across + tx * along;
        py = oy + ny * across + ty * along;
        if (jitter > 0.f) {
            const uint4 r = philox4x32_10(make_uint4((uint32_t)i, 0u, 0u, 0u),
                                          make_uint2((uint32_t)seed, (uint32_t)(seed >> 32)));
            const float u1 = ((float)(r.x >> 8) + 1.f) * (1.f / 16777216.f);  // (0, 1]
            const float u2 = (float)(r.y >> 8) * (1.f / 16777216.f);          // [0, 1)
            const float rad = sqrtf(-2.f * logf(u1));
            float sn, cs;
            sincosf(6.2831853071795864f * u2, &sn, &cs);
            px += rad * cs * jitter;
            py += rad * sn * jitter;
        }
        idx = i + count[1];
    }
    const bool contributor = active && idx < cap;
    const uint64_t mask = __ballot(contributor);
    if (contributor) {
        s.x[idx] = px;
        s.y[idx] = py;
        s.vx[idx] = velx * inv_tick;
        s.vy[idx] = vely * inv_tick;
        s.c00[idx] = 0.f;
        s.c01[idx] = 0.f;
        s.c10[idx] = 0.f;
        s.c11[idx] = 0.f;
        s.f00[idx] = 1.f;
        s.f01[idx] = 0.f;
        s.f10[idx] = 0.f;
        s.f11[idx] = 1.f;
        s.J[idx] = 1.f;
        s.mat[idx] = MPM_MAT_FLUID;
        s.id[idx] = idx;
        s.key[idx] = 0u;
    }
    if ((threadIdx.x & 63) == 0 && mask != 0ull) { atomicAdd(&count[0], (int)__popcll(mask)); }
}
__global__ void k_commit_count(int32_t* count) { count[1] = count[0]; }

hipError_t launch_emitter(hipStream_t st, const DevParams& p, const Store& s, int32_t* count, int cap, float ox,
                          float oy, float velx, float vely, int width, float spacing, float jitter, int n_emit,
                          uint64_t seed, float inv_tick)
{
    if (n_emit <= 0) { return hipSuccess; }
    hipLaunchKernelGGL(k_emitter, dim3((unsigned)((n_emit + kThreads - 1) / kThreads)), dim3(kThreads), 0, st, p, s,
                       count, cap, ox, oy, velx, vely, width, spacing, jitter, n_emit, seed, inv_tick);
    hipLaunchKernelGGL(k_commit_count, dim3(1), dim3(1), 0, st, count);
    return hipGetLastError();
}

// ===================================================================================== stress

// Stress diagnostic: kirchhoff_tau of each particle, the function every P2G kernel calls.  With
// snow_return set, snow particles first take the return mapping of G2P (snow_project on F as the trial
// elastic F and J as Jp, writing the projected F and Jp to F_out, J_out) and their stress then uses the
// mapping's rotation, as k_g2p2g does.
__global__ __launch_bounds__(kThreads) void k_stress(DevParams p, int n, const uint8_t* mat, const float* F,
                                                     const float* J, float* tau, int snow_return, float* F_out,
                                                     float* J_out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) { return; }
    Mat2 f = {F[i], F[n + i], F[2 * n + i], F[3 * n + i]};
    float j = J[i];
    const int m = mat[i];
    SnowHint hint;
    const SnowHint* rp = nullptr;
    if (snow_return && m == MPM_MAT_SNOW) {
        float rc, rs;
        snow_project(p, f, j, rc, rs);
        hint = {rc, rs, snow_hardening(p, j)};
        rp = &hint;
    }
    float t00, t01, t11;
    kirchhoff_tau(p, m, f, j, rp, t00, t01, t11);
    tau[i] = t00;
    tau[n + i] = t01;
    tau[2 * n + i] = t11;
    if (F_out) {
        F_out[i] = f.a00;
        F_out[n + i] = f.a01;
        F_out[2 * n + i] = f.a10;
        F_out[3 * n + i] = f.a11;
        J_out[i] = j;
    }
}

hipError_t launch_stress(hipStream_t st, const DevParams& p, int n, const uint8_t* mat, const float* F, const float* J,
                         float* tau, int snow_return, float* F_out, float* J_out)
{
    if (n <= 0) { return hipSuccess; }
    hipLaunchKernelGGL(k_stress, dim3((unsigned)((n + kThreads - 1) / kThreads)), dim3(kThreads), 0, st, p, n, mat, F,
                       J, tau, snow_return, F_out, J_out);
    return hipGetLastError();
}

#endif  // MPM_MAIN_TU

}  // namespace mpm
