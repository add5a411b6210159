// kernels.h — host-side launchers of the gfx950 kernels (defined in kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mpm_device.h"

namespace mpm {

// Output slot description read by the output kernel from a device symbol (the reference keeps it in a
// __constant__ View updated by cudaMemcpyToSymbolAsync before each tick, MPM.cu:13, 32-37).
struct OutView {
    float* pos;
    float* vel;
    int32_t* num_atoms;
    int32_t out_chunk;
    int32_t pad;
};

// Sort scratch of one handle.
struct SortScratch {
    uint32_t* keys_tmp;
    uint32_t* keys_alt;
    int32_t* vals_tmp;
    int32_t* vals_alt;
    int32_t* perm;
    int32_t* hist;       // kRadixBins * ntiles
    int32_t* partials;   // scan partials
    int32_t* bin_start;  // K
    int32_t* bin_end;    // K
    int32_t* item_cnt;   // K (scanned in place)
    Item* items;         // max_items
    int32_t* num_items;  // 1
    int capacity;
    int ntiles;
    int num_keys;
    int key_bits;
    int max_items;
    int chunk;
};

// Tile shapes compiled into the library (bin_x, bin_y); index used by the launchers.
constexpr int kNumShapes = 4;
constexpr int kShapeX[kNumShapes] = {8, 16, 16, 32};
constexpr int kShapeY[kNumShapes] = {8, 8, 16, 16};
// The binning's checkerboard slots need an even bin_x and bin_x * bin_y a power of two of at least 64.
constexpr bool shapes_fit_binning()
{
    for (int k = 0; k < kNumShapes; ++k) {
        const int nn = kShapeX[k] * kShapeY[k];
        if (kShapeX[k] % 2 != 0 || nn < 64 || (nn & (nn - 1)) != 0) { return false; }
    }
    return true;
}
static_assert(shapes_fit_binning(), "tile shape unsuited to the binning");
int shape_index(int bx, int by);

hipError_t launch_fill_f32(hipStream_t st, float* p, size_t n, float v);
hipError_t launch_fill_i32(hipStream_t st, int32_t* p, size_t n, int32_t v);

hipError_t launch_grid_update(hipStream_t st, const DevParams& p, const Grid& g);
hipError_t launch_p2g_reference(hipStream_t st, const DevParams& p, const Store& s, const int32_t* count, int cap,
                                const Grid& g);
hipError_t launch_g2p_reference(hipStream_t st, const DevParams& p, const Store& s, const int32_t* count, int cap,
                                const Grid& g);

// Stable binning: src (current order) -> dst (binned order), then bin ranges and work items.  full = false
// leaves v and C out of the gather (see k_gather).
hipError_t launch_binning(hipStream_t st, const DevParams& p, const Store& src, const Store& dst,
                          const int32_t* count, const SortScratch& ss, bool full = true);

hipError_t launch_p2g_tiled(hipStream_t st, int shape, const DevParams& p, const Store& s, const Grid& g,
                            const SortScratch& ss, const int32_t* count);
hipError_t launch_g2p_tiled(hipStream_t st, int shape, const DevParams& p, const Store& s, const Grid& g,
                            const SortScratch& ss, const int32_t* count);
hipError_t launch_g2p2g(hipStream_t st, int shape, const DevParams& p, const Store& s, const Grid& g,
                        const SortScratch& ss, const int32_t* count);
// The analytic-closure branch of launch_g2p2g (g2p2g.hip).
hipError_t launch_g2p2g_plain(hipStream_t st, int shape, const DevParams& p, const Store& s, const Grid& g,
                              const SortScratch& ss, const int32_t* count);

hipError_t launch_output(hipStream_t st, const Store& s, int32_t* count, int cap);
// Pack W2 of the closure (float[4675] parameters) into the MFMA A-operand layout (8 x 64 bf16x8).
size_t closure_packed_bytes();
hipError_t launch_pack_closure(hipStream_t st, const float* w, void* out);
hipError_t set_output_view_async(hipStream_t st, const OutView* host_pinned);

hipError_t launch_emitter(hipStream_t st, const DevParams& p, const Store& s, int32_t* count, int cap, float ox,
                          float oy, float velx, float vely, int width, float spacing, float jitter, int n_emit,
                          uint64_t seed, float inv_tick);

// Per-particle Kirchhoff stress (t00, t01, t11) of F (4 planes) and J for material mat[i]; with snow_return,
// snow particles take the G2P return mapping first (projected F and Jp into F_out, J_out, 4 planes and 1).
hipError_t launch_stress(hipStream_t st, const DevParams& p, int n, const uint8_t* mat, const float* F, const float* J,
                         float* tau, int snow_return, float* F_out, float* J_out);

// Deinterleave / interleave helpers for the boundary copies.
hipError_t launch_scatter_to_user(hipStream_t st, const Store& s, const int32_t* count, int cap, float* x2, float* v2,
                                  float* C4, float* F4, float* J, uint8_t* mat);

}  // namespace mpm
