// render.hip — headless particle rasteriser behind include/mpm_render.h.
//
// Reference counterparts (libs/render/source/cuda-render.cu): backgroundKernel (:214-243), drawKernel and its
// alpha blend (:244-310), launchFill (:189-213), cuda_engine_render_frame (:163-186).  The reference blends
// every particle into the image with an unsynchronised read-modify-write; here three passes make the same
// image deterministically:
//   1. k_render_clear     zero the per-pixel coverage counters
//   2. k_render_cover     one thread per particle adds 1 to each of the 21 pixels of its footprint
//   3. k_render_compose   background colour, then the 8-bit blend applied `coverage` times per channel
// Pass 3 uses per-channel tables of the blend p -> (uint8)(c * a + p * (1 - a) + 0.5) built on the host, so
// its result is exactly that of blending the particles one after another in any order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <mutex>
#include <string>

#include "../../include/mpm_render.h"

namespace {

constexpr int kThreads = 256;
constexpr int kColour[3] = {140, 210, 255};  // drawKernel's particle colour (cuda-render.cu:292-294)
constexpr int kAlpha = 100;                  // and alpha (cuda-render.cu:295)

struct BlendLut {
    uint8_t c[3][256];
};

struct RenderArgs {
    int w, h, pitch, chunk_px;
    float scale, flip_extent;
};

__global__ void k_render_clear(uint32_t* cov, size_t n)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        cov[i] = 0u;
    }
}

// Footprint of drawKernel (cuda-render.cu:281-308): pixel (int)(x * scale), (int)(y * scale) must lie in the
// image; then the 5x5 block around it without its four corners, clipped to the image.
__global__ void k_render_cover(RenderArgs a, const float* pos, int n, uint32_t* cov)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) { return; }
    const float x = pos[2 * i];
    float y = pos[2 * i + 1];
    if (a.flip_extent > 0.f) { y = a.flip_extent - y; }
    const float sx = x * a.scale, sy = y * a.scale;
    // (int) of a value outside int's range is undefined: such particles are off the image anyway
    if (!(sx > -2147483520.f && sx < 2147483520.f && sy > -2147483520.f && sy < 2147483520.f)) { return; }
    const int rx = (int)sx, ry = (int)sy;
    if ((unsigned)rx >= (unsigned)a.w || (unsigned)ry >= (unsigned)a.h) { return; }
    for (int di = -2; di <= 2; ++di) {
        const int py = ry + di;
        if (py < 0 || py >= a.h) { continue; }
        for (int dj = -2; dj <= 2; ++dj) {
            const int px = rx + dj;
            if (px < 0 || px >= a.w) { continue; }
            if ((di == 2 || di == -2) && (dj == 2 || dj == -2)) { continue; }
            atomicAdd(&cov[(size_t)py * a.w + px], 1u);
        }
    }
}

// backgroundKernel (cuda-render.cu:214-243) followed by `coverage` blends of the particle colour.
__global__ void k_render_compose(RenderArgs a, BlendLut lut, const uint32_t* cov, uint8_t* img)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= a.w) { return; }
    int c0, c1, c2;
    if (a.chunk_px > 0 && (x % a.chunk_px == 0 || y % a.chunk_px == 0)) {
        c0 = c1 = c2 = 0;
    } else {
        const int shade = (int)(uint8_t)(100.f * ((float)y / (float)a.h));
        c0 = (uint8_t)(255 - shade);
        c1 = (uint8_t)(160 - shade);
        c2 = 80;
    }
    const uint32_t k = cov[(size_t)y * a.w + x];
    for (uint32_t t = 0; t < k; ++t) {
        const int n0 = lut.c[0][c0], n1 = lut.c[1][c1], n2 = lut.c[2][c2];
        if (n0 == c0 && n1 == c1 && n2 == c2) { break; }  // fixed point: further blends change nothing
        c0 = n0;
        c1 = n1;
        c2 = n2;
    }
    uchar4* px = reinterpret_cast<uchar4*>(img + (size_t)y * a.pitch) + x;
    *px = make_uchar4((unsigned char)c0, (unsigned char)c1, (unsigned char)c2, 255);
}

thread_local int32_t g_render_error = 0;
thread_local std::string g_render_msg;

// The runtime code (hipError_t) is kept only when one is given, like record() in cuda-render.cu:23-27.
int32_t render_fail(int32_t code, int32_t detail, const std::string& msg)
{
    if (detail) { g_render_error = detail; }
    g_render_msg = msg;
    return code;
}

#define RENDER_TRY(call)                                                                               \
    do {                                                                                               \
        const hipError_t e_ = (call);                                                                  \
        if (e_ != hipSuccess) {                                                                        \
            return render_fail(MPM_ERR_DEVICE, (int32_t)e_, std::string(#call ": ") + hipGetErrorString(e_)); \
        }                                                                                              \
    } while (0)

// The blend of drawKernel's alpha() (cuda-render.cu:244-257) evaluated exactly as written, without
// contraction into fused multiply-adds.
#pragma clang fp contract(off)
BlendLut make_lut()
{
    BlendLut l;
    constexpr float inv255 = 1.f / 255.f;
    const float a = (float)kAlpha * inv255;
    const float ia = 1.f - a;
    for (int ch = 0; ch < 3; ++ch) {
        for (int p = 0; p < 256; ++p) {
            const float over = (float)kColour[ch] * a;
            const float under = (float)p * ia;
            l.c[ch][p] = (uint8_t)(over + under + 0.5f);
        }
    }
    return l;
}

// Checks a description; *pitch receives the row pitch in bytes.
int32_t check_desc(const MpmRenderDesc& d, int* pitch)
{
    if (d.width < 1 || d.height < 1 || d.width > 32768 || d.height > 32768) {
        return render_fail(MPM_ERR_INVALID_ARG, 0, "image size outside [1, 32768]");
    }
    *pitch = d.row_pitch ? d.row_pitch : 4 * d.width;
    if (*pitch < 4 * d.width || *pitch % 4 != 0) {
        return render_fail(MPM_ERR_INVALID_ARG, 0, "row_pitch below 4 * width or not a multiple of 4");
    }
    if (!(d.scale > 0.f) || !std::isfinite(d.scale) || !(d.flip_extent >= 0.f) || !std::isfinite(d.flip_extent) ||
        d.chunk_px < 0) {
        return render_fail(MPM_ERR_INVALID_ARG, 0, "scale must be > 0, flip_extent >= 0, chunk_px >= 0");
    }
    return MPM_OK;
}

int32_t check_view(const MpmView* view)
{
    if (view->count < 0 || (view->count > 0 && !view->pos)) {
        return render_fail(MPM_ERR_INVALID_ARG, 0, "view has a negative count or no positions");
    }
    return MPM_OK;
}

// The three passes of a frame on one stream: clear the counters of the image's pixels, count coverage,
// compose into target (device memory).
hipError_t enqueue_frame(hipStream_t st, const RenderArgs& a, const BlendLut& lut, uint32_t* cov, const MpmView& view,
                         uint8_t* target)
{
    const size_t px = (size_t)a.w * a.h;
    hipLaunchKernelGGL(k_render_clear, dim3((unsigned)std::min<size_t>((px + kThreads - 1) / kThreads, 65536)),
                       dim3(kThreads), 0, st, cov, px);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) { return e; }
    if (view.count > 0) {
        hipLaunchKernelGGL(k_render_cover, dim3((unsigned)((view.count + kThreads - 1) / kThreads)), dim3(kThreads), 0,
                           st, a, view.pos, view.count, cov);
        if ((e = hipGetLastError()) != hipSuccess) { return e; }
    }
    hipLaunchKernelGGL(k_render_compose, dim3((unsigned)((a.w + kThreads - 1) / kThreads), (unsigned)a.h),
                       dim3(kThreads), 0, st, a, lut, cov, target);
    return hipGetLastError();
}

// ---- the no-allocation frame (mpm_render_frame_noalloc) ----
// Coverage counters of mpm_render_frame_noalloc: one per pixel, for images of at most kScratchPixels pixels
// (8 MiB per device, held by every device the library runs on until the process ends).  Host code never
// uses the array's name as an address: scratch_on_device() looks the device address up once per device.
constexpr size_t kScratchPixels = (size_t)2048 * 1024;
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
uint32_t* g_scratch_ptr[kScratchDevices] = {};

}  // namespace

static __device__ __align__(256) uint32_t scratch[kScratchPixels];

namespace {

// Device address of `scratch` on device `dev`, which must be the current device (hipGetSymbolAddress
// answers for the current device; each device holds its own copy of the array).
hipError_t scratch_on_device(int dev, uint32_t** out)
{
    if (dev < 0 || dev >= kScratchDevices) { return hipErrorInvalidDevice; }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) { return e; }
        if (!p) { return hipErrorInvalidSymbol; }
        g_scratch_ptr[dev] = static_cast<uint32_t*>(p);
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

const BlendLut g_lut = make_lut();

}  // namespace

struct MpmRenderer {
    MpmRenderDesc d{};
    int pitch = 0;
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t* cov = nullptr;
    uint8_t* img = nullptr;  // device image for host targets
    BlendLut lut{};
};

extern "C" {

int32_t mpm_renderer_create(const MpmRenderDesc* desc, MpmRenderer** out)
{
    if (!desc || !out) { return render_fail(MPM_ERR_INVALID_ARG, 0, "desc or out is NULL"); }
    *out = nullptr;
    const MpmRenderDesc& d = *desc;
    int pitch = 0;
    const int32_t bad = check_desc(d, &pitch);
    if (bad != MPM_OK) { return bad; }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        return render_fail(MPM_ERR_NO_DEVICE, 0, "no HIP device");
    }
    MpmRenderer* r = new MpmRenderer();
    r->d = d;
    r->pitch = pitch;
    r->lut = make_lut();
    const size_t px = (size_t)d.width * d.height;
    hipError_t e = hipGetDevice(&r->device);
    if (e == hipSuccess) { e = hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking); }
    if (e == hipSuccess) { e = hipMalloc((void**)&r->cov, px * sizeof(uint32_t)); }
    if (e == hipSuccess) { e = hipMalloc((void**)&r->img, (size_t)pitch * d.height); }
    if (e != hipSuccess) {
        mpm_renderer_destroy(r);
        return render_fail(MPM_ERR_DEVICE, (int32_t)e, std::string("renderer allocation: ") + hipGetErrorString(e));
    }
    *out = r;
    return MPM_OK;
}

int32_t mpm_renderer_destroy(MpmRenderer* r)
{
    if (!r) { return MPM_ERR_INVALID_ARG; }
    if (r->cov) { (void)hipFree(r->cov); }
    if (r->img) { (void)hipFree(r->img); }
    if (r->stream) { (void)hipStreamDestroy(r->stream); }
    delete r;
    return MPM_OK;
}

int32_t mpm_render_frame(MpmRenderer* r, const MpmView* view, uint8_t* rgba, int32_t rgba_on_device)
{
    if (!r || !view || !rgba) { return render_fail(MPM_ERR_INVALID_ARG, 0, "renderer, view or rgba is NULL"); }
    if (check_view(view) != MPM_OK) { return MPM_ERR_INVALID_ARG; }
    const MpmRenderDesc& d = r->d;
    RenderArgs a{d.width, d.height, r->pitch, d.chunk_px, d.scale, d.flip_extent};
    uint8_t* target = rgba_on_device ? rgba : r->img;
    RENDER_TRY(hipSetDevice(r->device));
    RENDER_TRY(enqueue_frame(r->stream, a, r->lut, r->cov, *view, target));
    if (!rgba_on_device) {
        RENDER_TRY(hipMemcpyAsync(rgba, r->img, (size_t)r->pitch * d.height, hipMemcpyDeviceToHost, r->stream));
    }
    RENDER_TRY(hipStreamSynchronize(r->stream));
    return MPM_OK;
}

int32_t mpm_render_frame_noalloc(const MpmRenderDesc* desc, const MpmView* view, uint8_t* rgba_device)
{
    if (!desc || !view || !rgba_device) { return render_fail(MPM_ERR_INVALID_ARG, 0, "desc, view or rgba is NULL"); }
    int pitch = 0;
    const int32_t bad = check_desc(*desc, &pitch);
    if (bad != MPM_OK) { return bad; }
    if (check_view(view) != MPM_OK) { return MPM_ERR_INVALID_ARG; }
    if ((size_t)desc->width * (size_t)desc->height > kScratchPixels) {
        return render_fail(MPM_ERR_CAPACITY, 0,
                           "image of " + std::to_string(desc->width) + " x " + std::to_string(desc->height) +
                               " pixels exceeds the coverage scratch of " + std::to_string(kScratchPixels) +
                               " pixels of mpm_render_frame_noalloc");
    }
    int dev = 0;
    RENDER_TRY(hipGetDevice(&dev));
    uint32_t* cov = nullptr;
    RENDER_TRY(scratch_on_device(dev, &cov));
    const RenderArgs a{desc->width, desc->height, pitch, desc->chunk_px, desc->scale, desc->flip_extent};
    RENDER_TRY(enqueue_frame(nullptr, a, g_lut, cov, *view, rgba_device));
    RENDER_TRY(hipStreamSynchronize(nullptr));
    return MPM_OK;
}

int32_t mpm_write_ppm(const char* path, const uint8_t* rgba, int32_t width, int32_t height, int32_t row_pitch)
{
    if (!path || !rgba || width < 1 || height < 1) { return render_fail(MPM_ERR_INVALID_ARG, 0, "bad PPM arguments"); }
    const int pitch = row_pitch ? row_pitch : 4 * width;
    if (pitch < 4 * width) { return render_fail(MPM_ERR_INVALID_ARG, 0, "row_pitch below 4 * width"); }
    FILE* f = std::fopen(path, "wb");
    if (!f) { return render_fail(MPM_ERR_INVALID_ARG, 0, std::string("cannot open ") + path); }
    std::fprintf(f, "P6\n%d %d\n255\n", width, height);
    std::string row((size_t)width * 3, '\0');
    bool ok = true;
    for (int y = 0; y < height && ok; ++y) {
        const uint8_t* src = rgba + (size_t)y * pitch;
        for (int x = 0; x < width; ++x) {
            row[3 * x] = (char)src[4 * x];
            row[3 * x + 1] = (char)src[4 * x + 1];
            row[3 * x + 2] = (char)src[4 * x + 2];
        }
        ok = std::fwrite(row.data(), 1, row.size(), f) == row.size();
    }
    ok = (std::fclose(f) == 0) && ok;
    return ok ? MPM_OK : render_fail(MPM_ERR_INVALID_ARG, 0, std::string("write failed: ") + path);
}

int32_t mpm_render_last_error(void) { return g_render_error; }
const char* mpm_render_message(void) { return g_render_msg.c_str(); }

}  // extern "C"
