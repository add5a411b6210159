/*
 * mpm_engine.h — C-ABI of the MI355X MLS-MPM engine.
 *
 * This header replaces the C++ surface of the reference engine
 * (libs/cuda-engine/include/cuda-engine/cuda-engine.h) for the MPM path:
 *   - MPMArena         (cuda-engine.h:190-216, cuda-engine.cpp:146-174)  -> mpm_create / mpm_set_particles
 *   - GraphEnv         (cuda-engine.h:16-28)                             -> owned by the handle
 *   - instantiateMPMGraph (cuda-engine.h:227-230, MPM.cu:326-374)        -> mpm_create (graph built inside)
 *   - cudaGraphLaunch + t0/t1/tick_done events (main/MPM-master.cpp:106-109)
 *                                                                        -> mpm_step / mpm_query / mpm_sync /
 *                                                                           mpm_last_step_ms
 *   - OutputBuffer / View / MPMTickStaging (cuda-engine.h:51-80, 231-260; MPM.cu:19-44, 302-322)
 *                                                                        -> mpm_read_view (each tick rolls the
 *                                                                           slots on the device stream)
 *   - hosePipe         (cuda-engine.h:217-226, MPM.cu:439-473)            -> mpm_hose_pipe
 * and follows the error convention of the reference's only C-ABI (libs/render/include/render/cuda-interface.h:30-42,
 * libs/render/source/cuda-render.cu:23-27): int32_t status, 0 = ok, negative = error, plus last_error /
 * error_name queries.  No exception or abort crosses this boundary.
 *
 * All pointers are plain host pointers unless a comment says "device pointer".  Vector fields use the
 * reference's interleaved layout at the boundary: positions and velocities as float[2*n] (x0,y0,x1,y1,...)
 * like float2[], and C / F as float[4*n] row-major (F00,F01,F10,F11) like the reference's float4
 * (cuda-engine.h:199-204, cuda-internal.cuh:44-56).  Inside the engine the store is SoA per scalar.
 *
 * Threading: one host thread drives a handle (like the reference's sim thread).  mpm_read_view may be
 * called from a second thread (the reference's renderer thread reads OutputBuffer::readTarget,
 * cuda-engine.cpp:78-87).  Several handles may coexist; batched scenes live inside one handle.
 */
#ifndef MPM_ENGINE_H
#define MPM_ENGINE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ----
 * -1 .. -5 keep the meanings of the reference's only C-ABI (libs/render/source/cuda-render.cu):
 *   -1 a runtime call failed, its code in last_error (record(), :23-27);  -2 no matching device (:46);
 *   -3 not initialised / not bound (:80, :171);  -4 invalid argument (:31, :81, :89);
 *   -5 already initialised / bound (:32, :82-84).
 * -6 and -7 have no counterpart there.  MPM_BUSY is mpm_query's "still running" (cudaErrorNotReady,
 * main/MPM-master.cpp:83-84). */
#define MPM_OK 0
#define MPM_BUSY 1                  /* mpm_query: work still in flight                                      */
#define MPM_ERR_DEVICE (-1)         /* a HIP runtime call failed; mpm_last_error holds the hipError_t       */
#define MPM_ERR_NO_DEVICE (-2)      /* no usable gfx950 device                                               */
#define MPM_ERR_STATE (-3)          /* not ready: the call needs a state the handle is not in (no weights
                                       loaded, no tick run yet, no binning in reference mode)              */
#define MPM_ERR_INVALID_ARG (-4)    /* null handle/pointer, bad size, parameter out of range               */
#define MPM_ERR_ALREADY (-5)        /* reserved with the reference's meaning ("already initialised"); handles
                                       are created, never initialised twice, so no call returns it now      */
#define MPM_ERR_CAPACITY (-6)       /* particle count above the handle's capacity                            */
#define MPM_ERR_NOT_BUILT (-7)      /* requested code path is not compiled into this library                 */

/* ---- materials (per particle, uint8) ---- */
#define MPM_MAT_FLUID 0 /* Tait equation of state on the tracked J (reference MPM.cu:68-80)             */
#define MPM_MAT_JELLY 1 /* fixed-corotated elasticity on F (Stomakhin et al. 2012)                      */
#define MPM_MAT_SNOW 2  /* fixed-corotated + singular-value clamp plasticity (Stomakhin et al. 2013)   */

/* ---- execution modes ---- */
#define MPM_MODE_REFERENCE 0 /* reference kernel structure: per-particle P2G with HBM atomics, grid, G2P   */
#define MPM_MODE_TILED 1     /* binned particles, LDS-tiled P2G, grid, LDS-tiled G2P (three phases)         */
#define MPM_MODE_FUSED 2     /* binned particles, grid update + fused G2P->P2G (two phases per substep)     */

/* ---- constitutive closure ---- */
#define MPM_CLOSURE_ANALYTIC 0 /* polar decomposition + fixed-corotated / snow stress                     */
#define MPM_CLOSURE_LEARNED 1  /* BF16 MLP on (sigma1, sigma2, angle, J), analytic outside trained range */

/* ---- named parameter presets ---- */
#define MPM_PRESET_REFERENCE 0      /* reference constants.h: 196x100 nodes, dx=2, Tait fluid, 20 substeps */
#define MPM_PRESET_JELLY128 1       /* 128-cell unit square, dx=1/128, jelly E=1e3, dt=1e-4              */
#define MPM_PRESET_MIXED1024 2      /* 1024-cell unit square, fluid/jelly/snow boxes, dt=2e-5            */
#define MPM_PRESET_BATCH64X256 3    /* 64 scenes of 256 cells, mixed materials                           */
#define MPM_PRESET_LEARNED1024 4    /* MIXED1024 grid, jelly + snow, learned closure                     */
#define MPM_PRESET_LEARNED16X2048 5 /* 16 scenes of 2048 cells, jelly + snow, learned closure           */

/*
 * Simulation parameters.  Every compile-time constant of the reference (constants.h:9-88) is a field
 * here; mpm_params_preset(MPM_PRESET_REFERENCE) reproduces the reference values.
 */
typedef struct MpmParams {
    /* grid: nodes per scene including the halo (reference kGridDimX/Y = sim/dx + 2*kGridHalo, constants.h:76-77) */
    int32_t grid_w, grid_h;
    int32_t halo;      /* node offset of the domain origin, >= 2 (kGridHalo, constants.h:74)                  */
    int32_t wall;      /* node index <= wall or >= dim-1-wall zeroes the outward velocity (MPM.cu:159-168)   */
    float dx, inv_dx;  /* cell size and its reciprocal (kDx, inv_dx, constants.h:80-81)                       */
    float sim_w, sim_h;/* domain extent for the out-of-bounds clamp (kSimWidth/Height, MPM.cu:336-337)       */
    float dt;          /* substep (constants.h:88)                                                            */
    float gravity_x, gravity_y; /* added to node velocity times dt (kGravity on y, MPM.cu:154)              */
    float mass, vol;   /* particle mass and rest volume (kMass, kRestVol)                                     */
    float tait_k, tait_gamma, p_min, p_max; /* fluid: p = k((1/J)^gamma - 1) clamped (constants.h:84-87)     */
    float mu, lambda;  /* Lame parameters of jelly and snow                                                   */
    float theta_c, theta_s, hardening; /* snow: clamp [1-theta_c, 1+theta_s] (0 <= theta_c < 1, theta_s >= 0),
                                          mu,lambda *= exp(h(1-Jp))                                          */
    float j_min, j_max;/* clamp of the tracked J (MPM.cu:265)                                                 */
    int32_t substeps;  /* substeps per mpm_step tick (kSubsteps, constants.h:72)                              */
    int32_t capacity;  /* particle capacity over all scenes (kNumAtoms, constants.h:83)                       */
    int32_t num_scenes;/* independent scenes, each with its own grid_w x grid_h grid (batched form)           */
    int32_t sort_interval; /* substeps between re-binnings (>=1; binned modes only).  Up to `substeps`: within
                              a tick.  Above: a tick starts with a binning every ceil(sort_interval /
                              substeps) ticks, and always after the particles were set, loaded or emitted    */
    int32_t mode;      /* MPM_MODE_*                                                                          */
    int32_t closure;   /* MPM_CLOSURE_*                                                                       */
    int32_t block_x, block_y; /* cells per bin along x / y (0 = tuned default)                               */
    int32_t chunk;     /* particles per work item in the binned kernels (0 = tuned default)                   */
    int32_t out_chunk; /* particles per output chunk for View.num_atoms (kMaxAtoms = 1536, constants.h:50)   */
    int32_t reserved[7];
} MpmParams;

/* Device pointers of one output slot (reference View, cuda-engine.h:51-57). */
typedef struct MpmView {
    float* pos;         /* device pointer, float[2*capacity], particle i at pos[2i], pos[2i+1]             */
    float* vel;         /* device pointer, float[2*capacity]                                             */
    int32_t* num_atoms; /* device pointer, int32[ceil(capacity/out_chunk)] particles per output chunk      */
    int32_t slot;       /* 0..2, -1 before the first tick                                                  */
    int32_t count;      /* particles written to this slot                                                  */
} MpmView;

typedef struct MpmEngine MpmEngine;

/* Fill *out with a named preset (MPM_PRESET_*).  Replaces the constexpr block of constants.h:9-88. */
int32_t mpm_params_preset(int32_t preset, MpmParams* out);

/* Allocate the SoA store, grids and output triple buffer on the current device, and build the tick
 * hipGraph.  Replaces MPMArena() (cuda-engine.cpp:146-171), OutputBuffer (cuda-engine.h:62-80),
 * instantiateMPMGraph (MPM.cu:326-374) and MPMTickStaging::init (MPM.cu:19-30). */
int32_t mpm_create(const MpmParams* params, MpmEngine** out);
/* Release everything the handle owns (~MPMArena, ~GraphEnv, ~OutputArena: cuda-engine.cpp:172-174, 33-44, 63-66). */
int32_t mpm_destroy(MpmEngine* h);

/*
 * Replace the particle population (host arrays).  x, v: float[2n]; C, F: float[4n]; J: float[n];
 * material: uint8[n].  v, C, F, J, material may be NULL: v = 0, C = 0, F = I, J = 1 (the emitter's
 * initial state, MPM.cu:427-431), material = MPM_MAT_FLUID.  scene_counts: int32[num_scenes] particles
 * per scene, particles grouped by scene in order (NULL allowed when num_scenes == 1).  Particle i keeps
 * index i in every readback regardless of the engine's internal (binned) order.
 */
int32_t mpm_set_particles(MpmEngine* h, int32_t n, const float* x, const float* v, const float* C,
                          const float* F, const float* J, const uint8_t* material, const int32_t* scene_counts);

/* Enqueue n_ticks ticks of params.substeps substeps each, asynchronously (one graph launch per tick,
 * MPM-master.cpp:106-109).  Each tick ends with the output copy (MPMOutputKernel, MPM.cu:302-322) into the
 * write slot of the output triple buffer.  The write slot moves only when a finished tick is published
 * (mpm_query / mpm_sync, the reference's roll after tick_done, MPM-master.cpp:83-89), so every tick enqueued
 * between two publications writes the same slot, and no tick ever writes the published read slot or the
 * slot published before it (the cushion of cuda-engine.h:59-61). */
int32_t mpm_step(MpmEngine* h, int32_t n_ticks);
/* MPM_OK when all enqueued work finished, MPM_BUSY while in flight (cudaEventQuery, MPM-master.cpp:83-84).
 * On MPM_OK a tick written since the last publication is published: it becomes the read slot, and the
 * write slot moves to the slot after it (OutputBuffer::roll, cuda-engine.cpp:72-76). */
int32_t mpm_query(MpmEngine* h);
/* Block until all enqueued work finished, then publish like mpm_query. */
int32_t mpm_sync(MpmEngine* h);
/* GPU time of the last tick in ms (cudaEventElapsedTime(t0, t1), MPM-master.cpp:90-92). */
int32_t mpm_last_step_ms(MpmEngine* h, float* ms);

/* Device pointers of the slot most recently published (OutputBuffer::readTarget, cuda-engine.cpp:78-87).
 * Safe to call from a second thread.  The slot's contents stay unchanged until two more publications:
 * the next one moves the read slot on, and only the one after it lets a tick write this slot again, so a
 * reader has one publication period to finish with a view (the reference's third slot). */
int32_t mpm_read_view(MpmEngine* h, MpmView* out);

/* Synchronous host copy of the particle state in the caller's particle order.  Any pointer may be NULL.
 * Sizes as in mpm_set_particles; *n receives the current count. */
int32_t mpm_readback(MpmEngine* h, float* x, float* v, float* C, float* F, float* J, uint8_t* material,
                     int32_t* n);
/* Synchronous host copy of the grid of the last substep: mass and velocity per node, float[num_scenes*grid_w*grid_h]
 * each, node (scene, y, x) at scene*grid_w*grid_h + y*grid_w + x (reference node index, MPM.cu:122). */
int32_t mpm_read_grid(MpmEngine* h, float* mass, float* vel_x, float* vel_y);
/* Current particle count (device-resident count slot 0, cuda-engine.h:206-208). */
int32_t mpm_count(MpmEngine* h, int32_t* n);

/* ---- state file (checkpoint / resume; the reference has none) ----
 * Little-endian binary, a 200-byte header {char magic[8] = "MPMSTATE"; uint32 version = 1; uint32 header_bytes
 * = 200; int32 n; int32 num_scenes; uint64 checksum (64-bit FNV-1a of the payload); MpmParams params (164 bytes,
 * the saving handle's); uint32 reserved = 0}, then the payload in caller particle order: x f32[2n] |
 * v f32[2n] | C f32[4n] | F f32[4n] | J f32[n] | scene i32[n] | material u8[n]. */
/* Synchronous: write the current state (waits for enqueued work). */
int32_t mpm_save_state(MpmEngine* h, const char* path);
/* Replace the population with a saved one; particle i of the file becomes particle i of the handle.  The
 * file's grid size and scene count must equal the handle's (MPM_ERR_INVALID_ARG otherwise; also for a bad
 * magic, version, size or checksum); n above capacity gives MPM_ERR_CAPACITY.  The handle keeps its own
 * physical parameters. */
int32_t mpm_load_state(MpmEngine* h, const char* path);

/* Hose-pipe emitter (hosePipe, MPM.cu:439-473): enqueue width*rows particles on the handle's stream in
 * scene 0.  origin/velocity in scene units, velocity is the per-tick displacement (divided by the
 * tick length substeps*dt); jitter is the standard deviation of the Gaussian position jitter
 * (Philox4x32-10 + Box-Muller, counter = particle index, key = tick).  *emitted receives the count
 * enqueued; particles past capacity are dropped on the device like MPM.cu:422-436. */
int32_t mpm_hose_pipe(MpmEngine* h, float origin_x, float origin_y, float velocity_x, float velocity_y,
                      float nozzle_width, float spacing, float jitter, uint64_t tick, int32_t* emitted);

/* ---- learned constitutive closure (no counterpart in the reference, whose only stress is the Tait fluid,
 * MPM.cu:68-80) ----
 * A 4-64-64-3 SiLU network maps (s1, s2, theta, Jh) — singular values of the elastic F (s1 >= s2), angle of
 * U, and Jp for snow / 1 for jelly, all normalised — to tau / mu (tau00, tau01, tau11).  Parameters are a
 * float[MPM_CLOSURE_NUM_WEIGHTS] array: W1[64][4] | b1[64] | W2[64][64] | b2[64] | W3[3][64] | b3[3].
 * Inference runs as BF16 MFMA tiles with FP32 accumulation inside the binned G2P/P2G kernels; jelly and snow
 * particles outside s in [0.6, 1.6], Jh in [0.7, 1.3] and all fluid particles use the analytic stress. */
#define MPM_CLOSURE_NUM_WEIGHTS 4675

/* Training pairs from particle states: features float[4n] and targets float[3n] (analytic tau / mu) computed
 * on the device; valid[i] = 1 for jelly / snow particles inside the trained range. */
int32_t mpm_closure_samples(const MpmParams* params, int32_t n, const float* F, const float* J,
                            const uint8_t* material, float* features, float* targets, int32_t* valid);
/* Train on the GPU (hand-written forward, backward and Adam kernels, FP32): `steps` minibatches of `batch`
 * pairs drawn by a seeded hash from (features, targets), mean squared error.  weights: initial values in,
 * trained values out.  *final_loss: loss of the result on one more minibatch. */
int32_t mpm_closure_train(int32_t n, const float* features, const float* targets, int32_t steps, int32_t batch,
                          float lr, uint64_t seed, float* weights, float* final_loss);
/* One forward + backward over all n pairs on the GPU: mean loss and its gradient float[MPM_CLOSURE_NUM_WEIGHTS]. */
int32_t mpm_closure_grad(int32_t n, const float* features, const float* targets, const float* weights, float* grad,
                         float* loss);
/* Network outputs float[3n] of n feature vectors through the BF16 MFMA inference path of the engine. */
int32_t mpm_closure_eval(int32_t n, const float* features, const float* weights, float* out);
/* The inference path's SiLU: 256 linear segments (slope, intercept) as float[512], in index units
 * u = 16 z + 159.5 of the pre-activation z; segment clamp(rint(u), 0, 255) (round half to even) gives
 * silu(z) ~ slope u + intercept.
 * Host computation (no device needed); the same table is copied into every packed network. */
int32_t mpm_closure_act_table(float* slope_icpt);
/* Message of the last failed closure call. */
const char* mpm_closure_message(void);
/* Load trained parameters into a handle; required before mpm_step when params.closure is MPM_CLOSURE_LEARNED. */
int32_t mpm_set_closure(MpmEngine* h, const float* weights);

/* Autotuner result (no counterpart in the reference, whose launch shape is the compile-time 36 x 256 of
 * constants.h:18-21). */
typedef struct MpmTuneResult {
    int32_t block_x, block_y; /* bin / LDS tile shape                          */
    int32_t chunk;            /* particles per work item                       */
    int32_t sort_interval;    /* substeps between re-binnings                  */
    float ms_per_substep;     /* best measured GPU time per substep            */
    int32_t candidates;       /* configurations timed                          */
} MpmTuneResult;

/* Time candidate launch configurations of the binned modes on the caller's scene and return the fastest:
 * every compiled tile shape x chunk in {512, 1024, 2048, 4096, 8192} at base->sort_interval, then sort intervals
 * {base, 25, 50, 100} (<= substeps) for the winner.  Each candidate is a temporary handle that runs one
 * warm-up tick and one timed tick of `substeps` substeps (HIP events).  Arguments as mpm_set_particles.
 * No counterpart in the reference. */
int32_t mpm_autotune(const MpmParams* base, int32_t n, const float* x, const float* v, const uint8_t* material,
                     const int32_t* scene_counts, int32_t substeps, MpmTuneResult* out);

/* hipError_t of the last runtime call that failed on the handle, 0 if none (cuda_engine_last_error,
 * cuda-interface.h:41, cuda-render.cu:158); h may be NULL for failures before a handle existed (mpm_create,
 * mpm_autotune).  A status other than MPM_ERR_DEVICE leaves it unchanged. */
int32_t mpm_last_error(MpmEngine* h);
/* Name of a runtime error code as returned by mpm_last_error (cuda_engine_error_name: cudaGetErrorName,
 * cuda-render.cu:159). */
const char* mpm_error_name(int32_t err);
/* Static name of an MPM_* status code ("MPM_ERR_INVALID_ARG", ...; "MPM_UNKNOWN_STATUS" otherwise).  No
 * counterpart in the reference, whose statuses are bare integers. */
const char* mpm_status_name(int32_t status);
/* Human readable message of the last error on the handle (NULL handle: last create failure). */
const char* mpm_last_message(MpmEngine* h);

#ifdef __cplusplus
}
#endif
#endif /* MPM_ENGINE_H */
