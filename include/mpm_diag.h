/*
 * mpm_diag.h — diagnostic entry points of the MI355X MLS-MPM engine.
 *
 * None of these has a counterpart in the reference engine (libs/cuda-engine exposes no test or timing
 * hooks; its only instrument is the t0/t1 event pair around a whole tick, main/MPM-master.cpp:106-109).
 * They exist for the parity tests (bin order, stress) and for per-kernel timing with HIP events on the
 * engine's own stream.  All are synchronous and use the handle's stream.
 */
#ifndef MPM_DIAG_H
#define MPM_DIAG_H

#include <stdint.h>

#include "mpm_engine.h"

#ifdef __cplusplus
extern "C" {
#endif

/* kernel classes of mpm_time_kernels */
#define MPM_KT_BINNING 0 /* bin keys, radix passes, gather, bin ranges, work items (all launches of one binning) */
#define MPM_KT_P2G 1     /* k_p2g_reference or k_p2g_tiled                                                 */
#define MPM_KT_GRID 2    /* k_grid_update                                                                  */
#define MPM_KT_G2P 3     /* k_g2p_reference or k_g2p_tiled                                                 */
#define MPM_KT_G2P2G 4   /* k_g2p2g (fused G2P of substep s and P2G of substep s+1)                        */
#define MPM_KT_OUTPUT 5  /* k_output                                                                       */
#define MPM_KT_COUNT 6

/* Bin the current state exactly as a tick does first (stable radix sort by bin key and stencil base node
 * within the bin, then layers of one particle per node inside each bin, the nodes of a layer in checkerboard
 * order: lx + ly even first, then odd, each in node order) and copy the binned order out:
 * keys[j] = bin key of the j-th particle, ids[j] = its caller index.  Arrays sized for the current count;
 * *n receives it.  Not available in MPM_MODE_REFERENCE (returns MPM_ERR_STATE). */
int32_t mpm_debug_bin(MpmEngine* h, uint32_t* keys, int32_t* ids, int32_t* n);

/* Advance `substeps` substeps with the handle's mode, outside the tick graph: every run of consecutive
 * launches of one class (for example the fused substeps of a segment, which run back to back) is captured
 * into a graph of its own and launched between a pair of HIP events on the engine stream, so the kernels
 * run with a graph launch's dispatch gaps, as inside the tick graph.  ms[k] receives the
 * summed GPU time of class k, launches[k] the number of launches (arrays of MPM_KT_COUNT).  The output
 * slot is not written.  Advances the simulation exactly like mpm_step would: whole ticks follow mpm_step's
 * binning schedule (when sort_interval spans several ticks, only some ticks start with a binning); a
 * remainder shorter than a tick starts with one. */
int32_t mpm_time_kernels(MpmEngine* h, int32_t substeps, float* ms, int32_t* launches);

/* Kirchhoff stress of n particles with the handle's material parameters, computed by the device function
 * every P2G kernel calls (fluid: -J p I with the Tait pressure of MPM.cu:68-69; jelly and snow:
 * fixed-corotated with the polar rotation of F): F float[4n] row-major, J float[n] (Jp for snow), material
 * uint8[n]; tau float[3n] = (tau00, tau01, tau11) per particle. */
int32_t mpm_compute_stress(MpmEngine* h, int32_t n, const uint8_t* material, const float* F, const float* J,
                           float* tau);

/* The snow step of a fused substep on n particles: the return mapping of G2P on the trial elastic F
 * (F_trial float[4n], Jp float[n]) -> projected F_out float[4n] and Jp_out float[n], then the P2G stress
 * tau float[3n] of the projected state with the rotation the return mapping found, as the fused kernel
 * computes it. */
int32_t mpm_compute_snow_return(MpmEngine* h, int32_t n, const float* F_trial, const float* Jp, float* F_out,
                                float* Jp_out, float* tau);

#ifdef __cplusplus
}
#endif
#endif /* MPM_DIAG_H */
