/*
 * mpm_render.h — C-ABI of the headless particle rasteriser.
 *
 * Replaces the reference's render library for the MPM output (libs/render/include/render/cuda-interface.h,
 * libs/render/source/cuda-render.cu) without the D3D12 interop: the reference draws a View
 * (cuda-engine.h:51-57) into a swap-chain buffer shared through external memory and semaphores
 * (cuda-render.cu:29-156); here the target is a plain RGBA8 image in device or host memory.
 *   - cuda_engine_init / cuda_engine_bind_buffers (cuda-render.cu:29-156) -> mpm_renderer_create
 *   - cuda_engine_render_frame + launchFill   (cuda-render.cu:163-213)     -> mpm_render_frame
 *   - backgroundKernel (cuda-render.cu:214-243): vertical gradient, black chunk lines
 *   - drawKernel (cuda-render.cu:259-310): every particle blends colour (140, 210, 255) at alpha 100/255
 *     over the 5x5 pixel block around (int)(x * scale), (int)(y * scale), corners excluded
 * Status codes are those of mpm_engine.h (0 = ok, negative = error), with mpm_render_last_error /
 * mpm_render_message for the last failure on the calling thread.
 *
 * Determinism: the reference's drawKernel blends with unsynchronised read-modify-writes, so overlapping
 * particles race.  Every particle blends the same colour with the same alpha, so the serial result at a
 * pixel depends only on how many particles cover it: this rasteriser counts coverage with integer atomics
 * and applies the 8-bit blend that many times, which reproduces the serial reference loop bit for bit.
 */
#ifndef MPM_RENDER_H
#define MPM_RENDER_H

#include <stddef.h>
#include <stdint.h>

#include "mpm_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct MpmRenderDesc {
    int32_t width, height; /* image size in pixels (reference kDispWidth x kDispHeight = 1920 x 960)          */
    int32_t row_pitch;     /* bytes per image row, >= 4 * width (0 = 4 * width)                               */
    int32_t chunk_px;      /* period of the black grid lines in pixels (reference kChunkDim * kSimScale = 160);
                              0 = no lines                                                                    */
    float scale;           /* pixels per scene unit (reference kSimScale = 5)                                 */
    float flip_extent;     /* > 0: draw y as flip_extent - y (scenes with y up); 0: y down like the reference */
} MpmRenderDesc;

typedef struct MpmRenderer MpmRenderer;

/* Validate the description and allocate the per-pixel coverage buffer on the current device. */
int32_t mpm_renderer_create(const MpmRenderDesc* desc, MpmRenderer** out);
int32_t mpm_renderer_destroy(MpmRenderer* r);

/* Draw the first view->count particles of view->pos (device pointer, float[2 * count]) over the background
 * into rgba (RGBA8, desc.row_pitch bytes per row): a device pointer when rgba_on_device is nonzero, else host
 * memory (the image is drawn on the device and copied back).  Synchronous, like the reference's
 * cudaStreamSynchronize after launchFill (cuda-render.cu:179). */
int32_t mpm_render_frame(MpmRenderer* r, const MpmView* view, uint8_t* rgba, int32_t rgba_on_device);

/* No counterpart in the reference: mpm_render_frame for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  Draws the frame of *desc
 * (validated like mpm_renderer_create) into rgba_device, a device image of desc->row_pitch bytes per row,
 * with no renderer and no allocation: the per-pixel coverage counters live in a device array of the
 * library, 2048 x 1024 = 2,097,152 counters (8 MiB) per device, held by each device the library runs on
 * until the process ends.  An image of more pixels returns MPM_ERR_CAPACITY.  Runs on the current device's
 * null stream and returns when the frame is done.  Every call on one device shares that device's
 * counters: calls that overlap on one device (from several host threads) race and must be serialised by
 * the caller. */
int32_t mpm_render_frame_noalloc(const MpmRenderDesc* desc, const MpmView* view, uint8_t* rgba_device);

/* Write an RGBA8 host image as a binary PPM (P6, alpha dropped).  Host only. */
int32_t mpm_write_ppm(const char* path, const uint8_t* rgba, int32_t width, int32_t height, int32_t row_pitch);

/* hipError_t of the last runtime call that failed in a render call on this thread, 0 if none
 * (cuda_engine_last_error, cuda-render.cu:158), and the message of the last failed render call. */
int32_t mpm_render_last_error(void);
const char* mpm_render_message(void);

#ifdef __cplusplus
}
#endif
#endif /* MPM_RENDER_H */
