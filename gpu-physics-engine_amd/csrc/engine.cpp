// engine.cpp — host side of the MI355X MLS-MPM engine behind the C-ABI of include/mpm_engine.h.
//
// Reference counterparts (libs/cuda-engine):
//   MPMArena slab + carve (cuda-engine.cpp:146-174)         -> Engine::allocate (one hipMalloc, 256-B carve)
//   GraphEnv (cuda-engine.h:16-28, cuda-engine.cpp:33-44)    -> Engine stream / events / graph execs
//   instantiateMPMGraph (MPM.cu:326-374)                      -> Engine::build_graphs (stream capture)
//   OutputBuffer roll/readTarget/writeTarget (cuda-engine.cpp:72-98) -> write_slot / read_slot below
//   MPMTickStaging (MPM.cu:19-44)                             -> pinned OutView per slot + hipMemcpyToSymbolAsync
//   hosePipe (MPM.cu:439-473)                                 -> mpm_hose_pipe
#include <hip/hip_runtime.h>

#include <atomic>
#include <cmath>
#include <cstddef>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/mpm_engine.h"
#include "../../include/mpm_diag.h"
#include "kernels.h"

using namespace mpm;

namespace {
// How a tick starts its binning.  kBinAfterP2G: the tick's first P2G runs on the bins of the tick before
// (valid: the population and its order are unchanged since), then the binning gathers every component but v
// and C, which no kernel reads before the tick's last G2P writes them.  kBinFirst: a full binning before
// the first P2G (after an upload or emission, whose particles are in no bin, and in the non-fused modes).
enum BinStart { kBinNone = 0, kBinFirst = 1, kBinAfterP2G = 2 };
}  // namespace

struct MpmEngine {
    MpmParams params{};
    DevParams dp{};
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t t0 = nullptr, t1 = nullptr, done = nullptr;
    // Tick graphs by starting store c and by how the tick bins (BinStart): [c][kBinNone] reuses the bins of an
    // earlier tick (built only when sort_interval spans several ticks), [c][kBinFirst] bins before its first
    // P2G, [c][kBinAfterP2G] (fused mode) runs its first P2G on the bins of the tick before and bins after it.
    hipGraph_t graph[2][3] = {};
    hipGraphExec_t exec[2][3] = {};
    int flips[3] = {0, 0, 0};   // store swaps per tick (parity) of each variant
    int cur = 0;                // store holding the current state
    int ticks_per_sort = 1;     // ceil(sort_interval / substeps) in the binned modes
    int ticks_since_sort = 0;   // ticks run since the last tick that began with a binning
    bool need_sort = true;      // the population or its order changed outside a tick: the next tick bins
    uint8_t* slab = nullptr;
    size_t slab_bytes = 0;
    Store store[2]{};
    Grid grid{};
    // Fused mode: three sets of P2G accumulators (set 0 is grid.mom_x/mom_y/mass), all zero between segments.
    float* acc[3][3]{};
    SortScratch ss{};
    int32_t* count = nullptr;  // device [2]: slot 0 authoritative, slot 1 staged (cuda-engine.h:206-208)
    float* out_pos[3]{};
    float* out_vel[3]{};
    int32_t* out_num[3]{};
    OutView* staged = nullptr;  // pinned [3]
    std::atomic<int32_t> read_slot{-1};
    int32_t write_slot = 0;
    int32_t pending_slot = -1;
    int32_t last_count = 0;             // particle count as of the last enqueued work (host-tracked)
    std::atomic<int32_t> slot_count[3]{};  // particles written to each output slot
    int shape = 2;
    float* closure_w = nullptr;   // device float[4675]
    void* closure_w2 = nullptr;   // device network in the MFMA operand layout (about 15 KiB)
    bool closure_loaded = false;
    bool has_ticked = false;
    int32_t last_error = 0;
    std::string last_msg;
};

namespace {

thread_local int32_t g_create_error = 0;
thread_local std::string g_create_msg;

// Records a failure: the message always, the runtime code (hipError_t, nonzero) only when one is given, as the
// reference's record() keeps the runtime code for cuda_engine_last_error (cuda-render.cu:23-27).
int32_t fail(MpmEngine* h, int32_t code, int32_t detail, const std::string& msg)
{
    if (h) {
        if (detail) { h->last_error = detail; }
        h->last_msg = msg;
    } else {
        if (detail) { g_create_error = detail; }
        g_create_msg = msg;
    }
    return code;
}

#define HIP_TRY(h, call)                                                                                         \
    do {                                                                                                         \
        hipError_t e_ = (call);                                                                                  \
        if (e_ != hipSuccess) {                                                                                  \
            return fail((h), MPM_ERR_DEVICE, (int32_t)e_,                                                        \
                        std::string(#call) + ": " + hipGetErrorName(e_) + " (" + hipGetErrorString(e_) + ")");  \
        }                                                                                                        \
    } while (0)

size_t align_up(size_t x, size_t a) { return (x + a - 1) & ~(a - 1); }

struct Carver {
    uint8_t* base;
    size_t off = 0;
    template <class T>
    T* take(size_t count)
    {
        T* r = reinterpret_cast<T*>(base ? base + off : nullptr);
        off += align_up(count * sizeof(T), 256);
        return r;
    }
};

int ilog2_ceil(uint64_t v)
{
    int b = 0;
    while ((1ull << b) < v) { ++b; }
    return b;
}

void derive(const MpmParams& p, DevParams& d)
{
    d.W = p.grid_w;
    d.H = p.grid_h;
    d.halo = p.halo;
    d.wall = p.wall;
    d.G = p.grid_w * p.grid_h;
    d.num_scenes = p.num_scenes;
    d.bin_x = p.block_x;
    d.bin_y = p.block_y;
    d.nbx = (p.grid_w + p.block_x - 1) / p.block_x;
    d.nby = (p.grid_h + p.block_y - 1) / p.block_y;
    d.bps = d.nbx * d.nby;
    d.node_bits = ilog2_ceil((uint64_t)p.block_x * (uint64_t)p.block_y);
    d.dx = p.dx;
    d.inv_dx = p.inv_dx;
    d.sim_w = p.sim_w;
    d.sim_h = p.sim_h;
    d.dt = p.dt;
    d.gdt_x = p.gravity_x * p.dt;
    d.gdt_y = p.gravity_y * p.dt;
    d.mass = p.mass;
    d.vol = p.vol;
    d.tait_k = p.tait_k;
    d.tait_gamma = p.tait_gamma;
    d.p_min = p.p_min;
    d.p_max = p.p_max;
    d.mu = p.mu;
    d.lambda = p.lambda;
    d.theta_c = p.theta_c;
    d.theta_s = p.theta_s;
    d.hardening = p.hardening;
    d.j_min = p.j_min;
    d.j_max = p.j_max;
    d.closure = p.closure;
}

// Validates p (after defaults are filled); returns an error message or empty.
std::string validate(MpmParams& p)
{
    if (p.block_x == 0 || p.block_y == 0) {
        p.block_x = 16;
        p.block_y = 16;
    }
    if (p.chunk == 0) { p.chunk = 512; }
    if (p.sort_interval <= 0) { p.sort_interval = p.substeps; }
    if (p.out_chunk <= 0) { p.out_chunk = 1536; }
    if (p.grid_w < 8 || p.grid_h < 8) { return "grid_w and grid_h must be >= 8"; }
    if ((int64_t)p.grid_w * p.grid_h > (1ll << 30)) { return "grid too large"; }
    if (p.halo < 2) { return "halo must be >= 2 (quadratic stencil reach)"; }
    if (p.wall < 0 || 2 * p.wall + 2 >= p.grid_w || 2 * p.wall + 2 >= p.grid_h) { return "wall out of range"; }
    if (!(p.dx > 0.f) || !(p.inv_dx > 0.f) || !(p.dt > 0.f)) { return "dx, inv_dx and dt must be > 0"; }
    if (!(p.sim_w > 0.f) || !(p.sim_h > 0.f)) { return "sim_w and sim_h must be > 0"; }
    if ((double)p.sim_w * p.inv_dx + p.halo + 1.5 > (double)p.grid_w ||
        (double)p.sim_h * p.inv_dx + p.halo + 1.5 > (double)p.grid_h) {
        return "domain does not fit the grid: need sim * inv_dx + halo + 1.5 <= grid";
    }
    if (!(p.mass > 0.f) || !(p.vol > 0.f)) { return "mass and vol must be > 0"; }
    if (!(p.j_min > 0.f) || !(p.j_max >= p.j_min)) { return "J clamp range invalid"; }
    if (!(p.theta_c >= 0.f && p.theta_c < 1.f) || !(p.theta_s >= 0.f)) {
        return "snow thresholds: need 0 <= theta_c < 1 and theta_s >= 0";
    }
    if (p.substeps < 1 || p.substeps > 100000) { return "substeps out of range"; }
    if (p.capacity < 1 || p.capacity > 1000000000) { return "capacity out of range"; }
    if (p.num_scenes < 1 || p.num_scenes > kMaxScenes) { return "num_scenes out of range"; }
    if (p.mode < MPM_MODE_REFERENCE || p.mode > MPM_MODE_FUSED) { return "unknown mode"; }
    if (p.closure != MPM_CLOSURE_ANALYTIC && p.closure != MPM_CLOSURE_LEARNED) { return "unknown closure"; }
    if (shape_index(p.block_x, p.block_y) < 0) { return "block_x x block_y is not a compiled tile shape"; }
    if (p.chunk < 64 || p.chunk > 65536) { return "chunk out of range"; }
    const int64_t K = (int64_t)p.num_scenes * ((p.grid_w + p.block_x - 1) / p.block_x) *
                      ((p.grid_h + p.block_y - 1) / p.block_y);
    if (K * p.block_x * p.block_y > (1ll << 31)) { return "too many bins"; }
    return "";
}

int32_t allocate(MpmEngine* h)
{
    const MpmParams& p = h->params;
    const size_t cap = (size_t)p.capacity;
    const size_t nodes = (size_t)h->dp.G * (size_t)p.num_scenes;
    const int K = h->dp.bps * p.num_scenes;
    SortScratch& ss = h->ss;
    ss.capacity = p.capacity;
    ss.ntiles = (int)((cap + kRadixTile - 1) / kRadixTile);
    ss.num_keys = K;
    // sort key: bin (scene-major) above the stencil base node within the bin
    ss.key_bits = std::max(1, ilog2_ceil((uint64_t)K)) + h->dp.node_bits;
    ss.chunk = p.chunk;
    ss.max_items = (int)((cap + p.chunk - 1) / p.chunk) + K;
    const size_t scan_len = std::max((size_t)kRadixBins * ss.ntiles, (size_t)K);
    const size_t n_partials = scan_len / (kThreads * 16) + 2;
    const size_t out_chunks = (cap + p.out_chunk - 1) / p.out_chunk;

    for (int pass = 0; pass < 2; ++pass) {
        Carver c{pass == 0 ? nullptr : h->slab};
        for (int s = 0; s < 2; ++s) {
            Store& st = h->store[s];
            float** planes[13] = {&st.x, &st.y, &st.vx, &st.vy, &st.c00, &st.c01, &st.c10, &st.c11,
                                  &st.f00, &st.f01, &st.f10, &st.f11, &st.J};
            for (float** pl : planes) { *pl = c.take<float>(cap); }
            st.mat = c.take<uint8_t>(cap);
            st.id = c.take<int32_t>(cap);
            st.key = c.take<uint32_t>(cap);
        }
        h->grid.mom_x = c.take<float>(nodes);
        h->grid.mom_y = c.take<float>(nodes);
        h->grid.mass = c.take<float>(nodes);
        h->grid.vel = c.take<float2>(nodes);
        h->grid.vmass = c.take<float>(nodes);
        h->acc[0][0] = h->grid.mom_x;
        h->acc[0][1] = h->grid.mom_y;
        h->acc[0][2] = h->grid.mass;
        for (int b = 1; b < 3; ++b) {
            for (int k = 0; k < 3; ++k) { h->acc[b][k] = p.mode == MPM_MODE_FUSED ? c.take<float>(nodes) : nullptr; }
        }
        ss.keys_tmp = c.take<uint32_t>(cap);
        ss.keys_alt = c.take<uint32_t>(cap);
        ss.vals_tmp = c.take<int32_t>(cap);
        ss.vals_alt = c.take<int32_t>(cap);
        ss.perm = c.take<int32_t>(cap);
        ss.hist = c.take<int32_t>((size_t)kRadixBins * ss.ntiles);
        ss.partials = c.take<int32_t>(n_partials);
        ss.bin_start = c.take<int32_t>((size_t)K);
        ss.bin_end = c.take<int32_t>((size_t)K);
        ss.item_cnt = c.take<int32_t>((size_t)K);
        ss.items = c.take<Item>((size_t)ss.max_items);
        ss.num_items = c.take<int32_t>(1);
        h->count = c.take<int32_t>(2);
        h->closure_w = c.take<float>(4675);
        h->closure_w2 = c.take<uint8_t>(closure_packed_bytes());
        for (int k = 0; k < 3; ++k) {
            h->out_pos[k] = c.take<float>(2 * cap);
            h->out_vel[k] = c.take<float>(2 * cap);
            h->out_num[k] = c.take<int32_t>(out_chunks);
        }
        if (pass == 0) {
            h->slab_bytes = c.off;
            HIP_TRY(h, hipMalloc((void**)&h->slab, h->slab_bytes));
        }
    }
    // Zero the whole slab once (grid accumulators must start at zero; counts at zero).
    HIP_TRY(h, hipMemsetAsync(h->slab, 0, h->slab_bytes, h->stream));
    h->dp.cl.w = h->closure_w;
    h->dp.cl.packed = h->closure_w2;
    HIP_TRY(h, hipHostMalloc((void**)&h->staged, 3 * sizeof(OutView), hipHostMallocDefault));
    for (int k = 0; k < 3; ++k) {
        h->staged[k].pos = h->out_pos[k];
        h->staged[k].vel = h->out_vel[k];
        h->staged[k].num_atoms = h->out_num[k];
        h->staged[k].out_chunk = p.out_chunk;
        h->staged[k].pad = 0;
    }
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    return MPM_OK;
}

// Per-class timing of mpm_time_kernels; absent while the tick graphs are captured.  Consecutive launches of
// one class form a group (the fused substeps of a segment are one group).  Each group is captured from the
// stream into a graph of its own and launched between two HIP events, so its kernels run with the dispatch
// gaps of a graph launch, as in the tick graph, and no event packet sits between the kernels of a group.
// The events skip the system-scope release a default event performs (an L2 write-back that would
// otherwise be counted inside the bracketed kernels).
struct LaunchTimer {
    hipStream_t st = nullptr;
    std::vector<hipEvent_t> ev;  // one (start, end) pair per group
    std::vector<int> cls;        // class of each group
    std::vector<int> count;      // launches in each group
    std::vector<hipGraph_t> graphs;
    std::vector<hipGraphExec_t> execs;
    bool open = false;           // the last group is still being captured
    hipError_t begin(int k)
    {
        if (open && cls.back() == k) {
            ++count.back();
            return hipSuccess;
        }
        hipError_t e = close();
        if (e != hipSuccess) { return e; }
        hipEvent_t a, b;
        e = hipEventCreateWithFlags(&a, hipEventDisableSystemFence);
        if (e != hipSuccess) { return e; }
        e = hipEventCreateWithFlags(&b, hipEventDisableSystemFence);
        if (e != hipSuccess) {
            (void)hipEventDestroy(a);
            return e;
        }
        ev.push_back(a);
        ev.push_back(b);
        cls.push_back(k);
        count.push_back(1);
        e = hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal);
        open = e == hipSuccess;
        return e;
    }
    hipError_t end() { return hipSuccess; }  // the group ends when another class starts (or at close)
    // Ends the capture of the open group and launches it between its events.
    hipError_t close()
    {
        if (!open) { return hipSuccess; }
        open = false;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(st, &g);
        if (e != hipSuccess) { return e; }
        graphs.push_back(g);
        hipGraphExec_t x = nullptr;
        if ((e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0)) != hipSuccess) { return e; }
        execs.push_back(x);
        if ((e = hipGraphUpload(x, st)) != hipSuccess) { return e; }
        if ((e = hipEventRecord(ev[ev.size() - 2], st)) != hipSuccess) { return e; }
        if ((e = hipGraphLaunch(x, st)) != hipSuccess) { return e; }
        return hipEventRecord(ev.back(), st);
    }
    ~LaunchTimer()
    {
        if (open) {  // a failed enqueue left a capture open: end it so the stream is usable again
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(st, &g);
            if (g) { (void)hipGraphDestroy(g); }
        }
        for (hipGraphExec_t x : execs) { (void)hipGraphExecDestroy(x); }
        for (hipGraph_t g : graphs) { (void)hipGraphDestroy(g); }
        for (hipEvent_t e : ev) { (void)hipEventDestroy(e); }
    }
};

// Grid view of accumulator set `target` (scattered into), optionally computing velocities from set `src`
// and zeroing set `zero` (fused substep); -1 = none.
Grid grid_for(const MpmEngine* h, int target, int src, int zero)
{
    Grid g = h->grid;
    g.mom_x = h->acc[target][0];
    g.mom_y = h->acc[target][1];
    g.mass = h->acc[target][2];
    g.src_x = src >= 0 ? h->acc[src][0] : nullptr;
    g.src_y = src >= 0 ? h->acc[src][1] : nullptr;
    g.src_m = src >= 0 ? h->acc[src][2] : nullptr;
    g.zero_x = zero >= 0 ? h->acc[zero][0] : nullptr;
    g.zero_y = zero >= 0 ? h->acc[zero][1] : nullptr;
    g.zero_m = zero >= 0 ? h->acc[zero][2] : nullptr;
    g.zero_n = zero >= 0 ? (size_t)h->dp.G * (size_t)h->params.num_scenes : 0;
    return g;
}

// Enqueue `substeps` substeps on the stream starting from store c (c is updated to the store holding the
// result), then the output copy when `output` is set.  dirty: which accumulator sets hold data, before (in,
// all zero) and after (out); every segment checks it scatters only into zero sets.
hipError_t enqueue_substeps(MpmEngine* h, int& c, int substeps, bool output, LaunchTimer* tm, bool dirty[3],
                            BinStart bin = kBinFirst)
{
    const MpmParams& p = h->params;
    const DevParams& d = h->dp;
    hipStream_t st = h->stream;
    const int cap = p.capacity;
    hipError_t e = hipSuccess;
#define ENQ(k, x)                                                 \
    do {                                                          \
        if (tm && (e = tm->begin(k)) != hipSuccess) { return e; } \
        e = (x);                                                  \
        if (e != hipSuccess) { return e; }                        \
        if (tm && (e = tm->end()) != hipSuccess) { return e; }    \
    } while (0)
    if (p.mode == MPM_MODE_REFERENCE) {
        // Reference graph (MPM.cu:357-369): P2G -> grid -> G2P per substep; the reset is folded into the
        // grid update, which zeroes the accumulators after reading them.
        for (int s = 0; s < substeps; ++s) {
            ENQ(MPM_KT_P2G, launch_p2g_reference(st, d, h->store[c], h->count, cap, h->grid));
            ENQ(MPM_KT_GRID, launch_grid_update(st, d, h->grid));
            ENQ(MPM_KT_G2P, launch_g2p_reference(st, d, h->store[c], h->count, cap, h->grid));
        }
    } else {
        int done = 0;
        while (done < substeps) {
            const int L = std::min(p.sort_interval, substeps - done);
            const bool after_p2g = done == 0 && bin == kBinAfterP2G && p.mode == MPM_MODE_FUSED;
            if ((bin != kBinNone || done > 0) && !after_p2g) {
                ENQ(MPM_KT_BINNING, launch_binning(st, d, h->store[c], h->store[c ^ 1], h->count, h->ss));
                c ^= 1;
            }
            if (p.mode == MPM_MODE_TILED) {
                const Store& s = h->store[c];
                for (int k = 0; k < L; ++k) {
                    ENQ(MPM_KT_P2G, launch_p2g_tiled(st, h->shape, d, s, h->grid, h->ss, h->count));
                    ENQ(MPM_KT_GRID, launch_grid_update(st, d, h->grid));
                    ENQ(MPM_KT_G2P, launch_g2p_tiled(st, h->shape, d, s, h->grid, h->ss, h->count));
                }
            } else {
                // Fused: every segment starts and ends with all three accumulator sets zero.  P2G scatters
                // into set 0; substep k of the segment computes node velocities from set (k-1)%3, scatters
                // into set k%3 and zeroes set (k+1)%3 (read by the launch before, written by the launch
                // after); the closing grid update reads the last set, zeroes it and the one before.
                if (dirty[0] || dirty[1] || dirty[2]) { return hipErrorUnknown; }  // plan invariant
                int t = 0, prev = -1;
                ENQ(MPM_KT_P2G, launch_p2g_tiled(st, h->shape, d, h->store[c], grid_for(h, t, -1, -1), h->ss,
                                                 h->count));
                dirty[t] = true;
                if (after_p2g) {  // the grid sums do not depend on the particles' order
                    ENQ(MPM_KT_BINNING, launch_binning(st, d, h->store[c], h->store[c ^ 1], h->count, h->ss, false));
                    c ^= 1;
                }
                const Store& s = h->store[c];
                for (int k = 1; k < L; ++k) {
                    const int src = t;
                    t = k % 3;
                    const int z = (k + 1) % 3;
                    if (dirty[t]) { return hipErrorUnknown; }
                    ENQ(MPM_KT_G2P2G, launch_g2p2g(st, h->shape, d, s, grid_for(h, t, src, dirty[z] ? z : -1), h->ss,
                                                   h->count));
                    dirty[t] = true;
                    dirty[z] = false;
                    prev = src;
                }
                ENQ(MPM_KT_GRID, launch_grid_update(st, d, grid_for(h, t, -1, prev)));
                dirty[t] = false;
                if (prev >= 0) { dirty[prev] = false; }
                ENQ(MPM_KT_G2P, launch_g2p_tiled(st, h->shape, d, s, h->grid, h->ss, h->count));
            }
            done += L;
        }
    }
    if (output) { ENQ(MPM_KT_OUTPUT, launch_output(st, h->store[c], h->count, cap)); }
#undef ENQ
    return hipSuccess;
}

// How the next tick bins: not at all while the bins of an earlier tick serve (sort_interval spanning several
// ticks); after its first P2G when the bins of the tick before match the particles (fused mode); first
// otherwise: after the population changed outside a tick (set, load, emission), before any binning, in the
// non-fused modes.
BinStart next_tick_bins(const MpmEngine* h)
{
    const bool bins = h->ticks_per_sort <= 1 || h->need_sort || h->ticks_since_sort + 1 >= h->ticks_per_sort;
    if (!bins) { return kBinNone; }
    return h->need_sort || h->params.mode != MPM_MODE_FUSED ? kBinFirst : kBinAfterP2G;
}
void note_tick(MpmEngine* h, bool binned)
{
    if (binned) {
        h->ticks_since_sort = 0;
        h->need_sort = false;
    } else {
        ++h->ticks_since_sort;
    }
}

int32_t build_graphs(MpmEngine* h)
{
    const MpmParams& p = h->params;
    int segs = 0;
    h->ticks_per_sort = 1;
    if (p.mode != MPM_MODE_REFERENCE) {
        const int iv = std::min(p.sort_interval, p.substeps);
        segs = (p.substeps + iv - 1) / iv;
        h->ticks_per_sort = (p.sort_interval + p.substeps - 1) / p.substeps;
    }
    h->flips[kBinFirst] = segs & 1;
    h->flips[kBinAfterP2G] = segs & 1;
    h->flips[kBinNone] = (segs - 1) & 1;
    for (int b = kBinNone; b <= kBinAfterP2G; ++b) {
        if (b == kBinNone && h->ticks_per_sort <= 1) { continue; }
        if (b == kBinAfterP2G && p.mode != MPM_MODE_FUSED) { continue; }
        for (int c = 0; c < 2; ++c) {
            HIP_TRY(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
            int cc = c;
            bool dirty[3] = {false, false, false};
            hipError_t e = enqueue_substeps(h, cc, p.substeps, true, nullptr, dirty, (BinStart)b);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(h->stream, &g);
            if (e != hipSuccess) {
                if (g) { (void)hipGraphDestroy(g); }
                return fail(h, MPM_ERR_DEVICE, (int32_t)e, std::string("tick capture: ") + hipGetErrorName(e));
            }
            if (e2 != hipSuccess) { return fail(h, MPM_ERR_DEVICE, (int32_t)e2, "hipStreamEndCapture failed"); }
            h->graph[c][b] = g;
            HIP_TRY(h, hipGraphInstantiate(&h->exec[c][b], g, nullptr, nullptr, 0));
            // upload now: the first launch of an exec that is not uploaded leaves gaps between its kernels
            HIP_TRY(h, hipGraphUpload(h->exec[c][b], h->stream));
        }
    }
    return MPM_OK;
}

void release(MpmEngine* h)
{
    if (!h) { return; }
    if (h->stream) { (void)hipStreamSynchronize(h->stream); }
    for (int c = 0; c < 2; ++c) {
        for (int b = 0; b < 3; ++b) {
            if (h->exec[c][b]) { (void)hipGraphExecDestroy(h->exec[c][b]); }
            if (h->graph[c][b]) { (void)hipGraphDestroy(h->graph[c][b]); }
        }
    }
    if (h->t0) { (void)hipEventDestroy(h->t0); }
    if (h->t1) { (void)hipEventDestroy(h->t1); }
    if (h->done) { (void)hipEventDestroy(h->done); }
    if (h->slab) { (void)hipFree(h->slab); }
    if (h->staged) { (void)hipHostFree(h->staged); }
    if (h->stream) { (void)hipStreamDestroy(h->stream); }
}

float f32_ref_dt() { return 0.032f / 20.f; }

// Upload a population; scene[i] is the scene of particle i (any order).  Particle i keeps id i.
int32_t set_particles_impl(MpmEngine* h, int32_t n, const float* x, const float* v, const float* C, const float* F,
                           const float* J, const uint8_t* material, const uint32_t* scene)
{
    const MpmParams& p = h->params;
    std::vector<uint32_t> scene_key((size_t)n, 0u);
    for (int32_t i = 0; i < n; ++i) {
        if (scene[i] >= (uint32_t)p.num_scenes) { return fail(h, MPM_ERR_INVALID_ARG, 0, "scene index out of range"); }
        scene_key[(size_t)i] = scene[i] * (uint32_t)h->dp.bps;
    }
    for (int32_t i = 0; i < n; ++i) {
        const float px = x[2 * i], py = x[2 * i + 1];
        if (!(px >= 0.f && px <= p.sim_w && py >= 0.f && py <= p.sim_h)) {
            return fail(h, MPM_ERR_INVALID_ARG, 0, "particle position outside [0,sim_w]x[0,sim_h] or not finite");
        }
        if (material && material[i] > MPM_MAT_SNOW) { return fail(h, MPM_ERR_INVALID_ARG, 0, "unknown material"); }
    }
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    const size_t N = (size_t)n;
    std::vector<float> planes(13 * N);
    for (size_t i = 0; i < N; ++i) {
        planes[0 * N + i] = x[2 * i];
        planes[1 * N + i] = x[2 * i + 1];
        planes[2 * N + i] = v ? v[2 * i] : 0.f;
        planes[3 * N + i] = v ? v[2 * i + 1] : 0.f;
        for (int k = 0; k < 4; ++k) {
            planes[(4 + k) * N + i] = C ? C[4 * i + k] : 0.f;
            planes[(8 + k) * N + i] = F ? F[4 * i + k] : ((k == 0 || k == 3) ? 1.f : 0.f);
        }
        planes[12 * N + i] = J ? J[i] : 1.f;
    }
    std::vector<uint8_t> mat(N);
    std::vector<int32_t> ids(N);
    for (size_t i = 0; i < N; ++i) {
        mat[i] = material ? material[i] : (uint8_t)MPM_MAT_FLUID;
        ids[i] = (int32_t)i;
    }
    Store& s = h->store[h->cur];
    float* dst[13] = {s.x, s.y, s.vx, s.vy, s.c00, s.c01, s.c10, s.c11, s.f00, s.f01, s.f10, s.f11, s.J};
    if (n > 0) {
        for (int k = 0; k < 13; ++k) {
            HIP_TRY(h, hipMemcpy(dst[k], planes.data() + (size_t)k * N, N * sizeof(float), hipMemcpyHostToDevice));
        }
        HIP_TRY(h, hipMemcpy(s.mat, mat.data(), N, hipMemcpyHostToDevice));
        HIP_TRY(h, hipMemcpy(s.id, ids.data(), N * sizeof(int32_t), hipMemcpyHostToDevice));
        HIP_TRY(h, hipMemcpy(s.key, scene_key.data(), N * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    const int32_t cnt[2] = {n, n};
    HIP_TRY(h, hipMemcpy(h->count, cnt, sizeof(cnt), hipMemcpyHostToDevice));
    h->last_count = n;
    h->need_sort = true;
    return MPM_OK;
}



// ---- state file (mpm_save_state / mpm_load_state) ----
// Little-endian: StateHeader, then the payload x f32[2n] | v f32[2n] | C f32[4n] | F f32[4n] | J f32[n] |
// scene i32[n] | material u8[n], checksummed with 64-bit FNV-1a.
constexpr char kStateMagic[8] = {'M', 'P', 'M', 'S', 'T', 'A', 'T', 'E'};
constexpr uint32_t kStateVersion = 1;

struct StateHeader {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    int32_t n;
    int32_t num_scenes;
    uint64_t checksum;
    MpmParams params;
    uint32_t reserved;  // zero; pads the header to a multiple of 8 bytes
};
static_assert(sizeof(MpmParams) == 164 && sizeof(StateHeader) == 200, "state header layout");
static_assert(offsetof(StateHeader, params) == 32, "state header layout");

uint64_t fnv1a64(const uint8_t* b, size_t n)
{
    uint64_t hsh = 0xcbf29ce484222325ull;  // FNV-1a 64-bit offset basis
    for (size_t i = 0; i < n; ++i) {
        hsh ^= b[i];
        hsh *= 0x100000001b3ull;  // FNV 64-bit prime
    }
    return hsh;
}

struct StatePayload {
    std::vector<uint8_t> bytes;
    float *x, *v, *C, *F, *J;
    uint8_t *scene, *mat;
    explicit StatePayload(size_t n) : bytes(n * 57)
    {
        float* f = reinterpret_cast<float*>(bytes.data());
        x = f;
        v = x + 2 * n;
        C = v + 2 * n;
        F = C + 4 * n;
        J = F + 4 * n;
        scene = reinterpret_cast<uint8_t*>(J + n);
        mat = scene + 4 * n;
    }
};

}  // namespace

extern "C" {

int32_t mpm_params_preset(int32_t preset, MpmParams* out)
{
    if (!out) { return MPM_ERR_INVALID_ARG; }
    MpmParams p;
    std::memset(&p, 0, sizeof(p));
    // Fields common to the unit-square presets.
    // rho: particle density.  With fixed-corotated jelly E = 1e4, nu = 0.2 the elastic wave speed is
    // sqrt((lambda + 2 mu) / rho); rho = 16 keeps c dt / dx at about 0.54 on the 1024-cell grid at
    // dt = 2e-5 (rho = 1 diverges within 10 substeps; rho = 4, CFL 1.08, amplifies rounding noise over
    // several hundred substeps).
    auto unit_square = [&](int cells, float dt, float rho) {
        p.grid_w = p.grid_h = cells + 4;
        p.halo = 2;
        p.wall = 2;
        p.dx = 1.f / (float)cells;
        p.inv_dx = (float)cells;
        p.sim_w = p.sim_h = 1.f;
        p.dt = dt;
        p.gravity_x = 0.f;
        p.gravity_y = -9.8f;
        p.vol = p.dx * p.dx / 4.f;
        p.mass = rho * p.vol;
        // fluid: bulk modulus 400 as Tait with gamma = 1 (tau = k (J - 1) I), no tension (as the reference)
        p.tait_k = 400.f;
        p.tait_gamma = 1.f;
        p.p_min = -0.f;
        p.p_max = 4000.f;
        p.theta_c = 2.5e-2f;
        p.theta_s = 7.5e-3f;
        p.hardening = 10.f;
        p.j_min = 0.1f;
        p.j_max = 10.f;
        p.num_scenes = 1;
        p.mode = MPM_MODE_FUSED;
        p.closure = MPM_CLOSURE_ANALYTIC;
        // launch shape chosen by mpm_autotune on MI355X for the 4,194,304-particle mixed scene
        p.block_x = 16;
        p.block_y = 8;
        p.chunk = 2048;
        p.out_chunk = 1536;
    };
    auto lame = [&](float E, float nu) {
        p.mu = E / (2.f * (1.f + nu));
        p.lambda = E * nu / ((1.f + nu) * (1.f - 2.f * nu));
    };
    switch (preset) {
    case MPM_PRESET_REFERENCE:
        // constants.h:9-88 and MPM.cu:336-337 (sim 384 x 192, grid 196 x 100, dx 2, halo 2)
        p.grid_w = 384 / 2 + 4;
        p.grid_h = 192 / 2 + 4;
        p.halo = 2;
        p.wall = 2;
        p.dx = 2.f;
        p.inv_dx = 1.f / 2.f;
        p.sim_w = 384.f;
        p.sim_h = 192.f;
        p.dt = f32_ref_dt();
        p.gravity_x = 0.f;
        p.gravity_y = 700.f;
        p.mass = 1.f;
        p.vol = 1.2f;
        p.tait_k = 8.2e4f;
        p.tait_gamma = 3.2f;
        p.p_min = -0.f * 8.2e4f;
        p.p_max = 10.f * 8.2e4f;
        p.j_min = 0.1f;
        p.j_max = 10.f;
        p.substeps = 20;
        p.capacity = 400 * 1000;
        p.num_scenes = 1;
        p.sort_interval = 20;
        p.mode = MPM_MODE_FUSED;
        p.closure = MPM_CLOSURE_ANALYTIC;
        p.block_x = 16;
        p.block_y = 16;
        p.chunk = 512;
        p.out_chunk = 128 * 12;
        break;
    case MPM_PRESET_JELLY128:
        unit_square(128, 1e-4f, 1.f);
        lame(1e3f, 0.2f);
        p.substeps = 100;
        p.capacity = 32768;
        p.sort_interval = 50;
        break;
    case MPM_PRESET_MIXED1024:
        unit_square(1024, 2e-5f, 16.f);
        lame(1e4f, 0.2f);
        p.substeps = 100;
        p.capacity = 4194304;
        p.sort_interval = 100;  // every tick: as fast as every second tick in ticks 6-25, and it stays at
                                // 7.5-8.0 ms per tick to tick 60, where every second tick reaches 26.6 ms
        // 16 x 16 bins of about 2,000 particles, one work item each (mpm_autotune's choice): 6.95-6.99 ms per
        // tick against 7.23-7.30 ms with 16 x 8 / 2048, and 7.56 against 8.17 ms at tick 50
        p.block_x = 16;
        p.block_y = 16;
        p.chunk = 4096;
        break;
    case MPM_PRESET_BATCH64X256:
        unit_square(256, 2e-5f, 16.f);
        lame(1e4f, 0.2f);
        p.substeps = 200;
        p.num_scenes = 64;
        p.capacity = 64 * 524288;
        p.sort_interval = 200;  // one binning per tick: 29.3 ms per tick against 30.6 sorting every 100
        p.chunk = 3072;         // steady-state sweep: 2304..4096 all ~1.5 % faster than 2048
        break;
    case MPM_PRESET_LEARNED1024:
        unit_square(1024, 2e-5f, 16.f);
        lame(1e4f, 0.2f);
        p.substeps = 100;
        p.capacity = 4194304;
        p.sort_interval = 100;
        p.closure = MPM_CLOSURE_LEARNED;
        break;
    case MPM_PRESET_LEARNED16X2048:
        unit_square(2048, 1e-5f, 16.f);
        lame(1e4f, 0.2f);
        p.substeps = 100;
        p.num_scenes = 16;
        p.capacity = 16 * 16777216 / 2;
        p.sort_interval = 100;
        // 16 x 16 bins of about 4,000 particles (16 per cell), one work item each: 390-394 ms per tick against
        // 403-404 ms with 16 x 8 / 3072 (the analytic engine on the same scenes: 156.5-157.1 against 161-162)
        p.block_x = 16;
        p.block_y = 16;
        p.chunk = 8192;
        p.closure = MPM_CLOSURE_LEARNED;
        break;
    default:
        return MPM_ERR_INVALID_ARG;
    }
    *out = p;
    return MPM_OK;
}

int32_t mpm_create(const MpmParams* params, MpmEngine** out)
{
    if (!params || !out) { return fail(nullptr, MPM_ERR_INVALID_ARG, 0, "null argument"); }
    *out = nullptr;
    MpmParams p = *params;
    const std::string bad = validate(p);
    if (!bad.empty()) { return fail(nullptr, MPM_ERR_INVALID_ARG, 0, bad); }
    if (p.closure == MPM_CLOSURE_LEARNED && p.mode == MPM_MODE_REFERENCE) {
        return fail(nullptr, MPM_ERR_NOT_BUILT, 0, "the learned closure runs in the binned modes only");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        return fail(nullptr, MPM_ERR_NO_DEVICE, 0, "no HIP device");
    }
    MpmEngine* h = new (std::nothrow) MpmEngine();
    if (!h) { return fail(nullptr, MPM_ERR_DEVICE, (int32_t)hipErrorOutOfMemory, "out of host memory"); }
    h->params = p;
    derive(p, h->dp);
    h->shape = shape_index(p.block_x, p.block_y);
    int32_t rc = MPM_OK;
    auto bail = [&](int32_t code) {
        g_create_error = h->last_error;
        g_create_msg = h->last_msg;
        release(h);
        delete h;
        return code;
    };
    hipError_t e = hipGetDevice(&h->device);
    if (e != hipSuccess) { return bail(fail(h, MPM_ERR_DEVICE, (int32_t)e, "hipGetDevice")); }
    e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    if (e != hipSuccess) { return bail(fail(h, MPM_ERR_DEVICE, (int32_t)e, "hipStreamCreateWithFlags")); }
    if ((e = hipEventCreate(&h->t0)) != hipSuccess || (e = hipEventCreate(&h->t1)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&h->done, hipEventDisableTiming)) != hipSuccess) {
        return bail(fail(h, MPM_ERR_DEVICE, (int32_t)e, "hipEventCreate"));
    }
    if ((rc = allocate(h)) != MPM_OK) { return bail(rc); }
    if ((rc = build_graphs(h)) != MPM_OK) { return bail(rc); }
    *out = h;
    return MPM_OK;
}

int32_t mpm_destroy(MpmEngine* h)
{
    if (!h) { return MPM_ERR_INVALID_ARG; }
    release(h);
    delete h;
    return MPM_OK;
}

int32_t mpm_set_particles(MpmEngine* h, int32_t n, const float* x, const float* v, const float* C, const float* F,
                          const float* J, const uint8_t* material, const int32_t* scene_counts)
{
    if (!h) { return MPM_ERR_INVALID_ARG; }
    const MpmParams& p = h->params;
    if (n < 0 || (n > 0 && !x)) { return fail(h, MPM_ERR_INVALID_ARG, 0, "n < 0 or x is NULL"); }
    if (n > p.capacity) { return fail(h, MPM_ERR_CAPACITY, 0, "n exceeds capacity"); }
    if (p.num_scenes > 1 && !scene_counts) { return fail(h, MPM_ERR_INVALID_ARG, 0, "scene_counts required"); }
    std::vector<uint32_t> scene((size_t)n, 0u);
    if (scene_counts) {
        int64_t total = 0;
        for (int s = 0; s < p.num_scenes; ++s) {
            if (scene_counts[s] < 0) { return fail(h, MPM_ERR_INVALID_ARG, 0, "negative scene count"); }
            for (int32_t k = 0; k < scene_counts[s]; ++k) {
                if (total + k < n) { scene[(size_t)(total + k)] = (uint32_t)s; }
            }
            total += scene_counts[s];
        }
        if (total != n) { return fail(h, MPM_ERR_INVALID_ARG, 0, "scene_counts do not sum to n"); }
    }
    return set_particles_impl(h, n, x, v, C, F, J, material, scene.data());
}


int32_t mpm_step(MpmEngine* h, int32_t n_ticks)
{
    if (!h) { return MPM_ERR_INVALID_ARG; }
    if (n_ticks < 0) { return fail(h, MPM_ERR_INVALID_ARG, 0, "n_ticks < 0"); }
    if (h->params.closure == MPM_CLOSURE_LEARNED && !h->closure_loaded) {
        return fail(h, MPM_ERR_STATE, 0, "learned closure selected but no weights loaded (mpm_set_closure)");
    }
    if (n_ticks == 0) { return MPM_OK; }
    // Every tick until the next publication writes the same slot: the write slot moves only in publish().
    const int32_t slot = h->write_slot;
    HIP_TRY(h, set_output_view_async(h->stream, &h->staged[slot]));
    for (int32_t k = 0; k < n_ticks; ++k) {
        const bool last = k == n_ticks - 1;
        const BinStart bin = next_tick_bins(h);
        if (last) { HIP_TRY(h, hipEventRecord(h->t0, h->stream)); }
        HIP_TRY(h, hipGraphLaunch(h->exec[h->cur][bin], h->stream));
        if (last) { HIP_TRY(h, hipEventRecord(h->t1, h->stream)); }
        h->cur ^= h->flips[bin];
        note_tick(h, bin != kBinNone);
        h->has_ticked = true;
    }
    h->slot_count[slot].store(h->last_count, std::memory_order_relaxed);
    h->pending_slot = slot;
    HIP_TRY(h, hipEventRecord(h->done, h->stream));
    return MPM_OK;
}

namespace {
// OutputBuffer::roll (cuda-engine.cpp:72-76), called once the ticks writing the pending slot have finished:
// that slot becomes the read slot and the write slot moves to the one after it.  Slots then go round in
// order, so the new write slot is neither the read slot nor the slot read before it (the cushion for a
// reader that took the previous read slot just before this store).
void publish(MpmEngine* h)
{
    if (h->pending_slot < 0) { return; }
    const int32_t s = h->pending_slot;
    h->read_slot.store(s, std::memory_order_release);
    h->write_slot = (s + 1) % 3;
    h->pending_slot = -1;
}
}  // namespace

int32_t mpm_query(MpmEngine* h)
{
    if (!h) { return MPM_ERR_INVALID_ARG; }
    const hipError_t e = hipEventQuery(h->done);
    if (e == hipErrorNotReady) { return MPM_BUSY; }
    if (e != hipSuccess) { return fail(h, MPM_ERR_DEVICE, (int32_t)e, "hipEventQuery"); }
    publish(h);
    return MPM_OK;
}

int32_t mpm_sync(MpmEngine* h)
{
    if (!h) { return MPM_ERR_INVALID_ARG; }
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    publish(h);
    return MPM_OK;
}

int32_t mpm_last_step_ms(MpmEngine* h, float* ms)
{
    if (!h || !ms) { return MPM_ERR_INVALID_ARG; }
    if (!h->has_ticked) { return fail(h, MPM_ERR_STATE, 0, "no tick has run"); }
    HIP_TRY(h, hipEventSynchronize(h->t1));
    HIP_TRY(h, hipEventElapsedTime(ms, h->t0, h->t1));
    return MPM_OK;
}

int32_t mpm_read_view(MpmEngine* h, MpmView* out)
{
    if (!h || !out) { return MPM_ERR_INVALID_ARG; }
    const int32_t s = h->read_slot.load(std::memory_order_acquire);
    out->slot = s;
    if (s < 0) {
        out->pos = nullptr;
        out->vel = nullptr;
        out->num_atoms = nullptr;
        out->count = 0;
        return MPM_OK;
    }
    out->pos = h->out_pos[s];
    out->vel = h->out_vel[s];
    out->num_atoms = h->out_num[s];
    out->count = h->slot_count[s].load(std::memory_order_relaxed);
    return MPM_OK;
}

int32_t mpm_count(MpmEngine* h, int32_t* n)
{
    if (!h || !n) { return MPM_ERR_INVALID_ARG; }
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    int32_t c[2];
    HIP_TRY(h, hipMemcpy(c, h->count, sizeof(c), hipMemcpyDeviceToHost));
    *n = c[0];
    h->last_count = c[0];
    return MPM_OK;
}

int32_t mpm_readback(MpmEngine* h, float* x, float* v, float* C, float* F, float* J, uint8_t* material, int32_t* n)
{
    if (!h) { return MPM_ERR_INVALID_ARG; }
    int32_t cnt = 0;
    int32_t rc = mpm_count(h, &cnt);
    if (rc != MPM_OK) { return rc; }
    if (n) { *n = cnt; }
    const size_t N = (size_t)cnt;
    if (N == 0) { return MPM_OK; }
    const Store& s = h->store[h->cur];
    std::vector<int32_t> ids(N);
    HIP_TRY(h, hipMemcpy(ids.data(), s.id, N * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (size_t j = 0; j < N; ++j) {
        if (ids[j] < 0 || (size_t)ids[j] >= N) { return fail(h, MPM_ERR_STATE, 0, "particle ids corrupted"); }
    }
    std::vector<float> tmp(N);
    auto fetch = [&](const float* src, float* dst, int stride, int k) -> int32_t {
        HIP_TRY(h, hipMemcpy(tmp.data(), src, N * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t j = 0; j < N; ++j) { dst[(size_t)ids[j] * stride + k] = tmp[j]; }
        return MPM_OK;
    };
    if (x && ((rc = fetch(s.x, x, 2, 0)) || (rc = fetch(s.y, x, 2, 1)))) { return rc; }
    if (v && ((rc = fetch(s.vx, v, 2, 0)) || (rc = fetch(s.vy, v, 2, 1)))) { return rc; }
    if (C && ((rc = fetch(s.c00, C, 4, 0)) || (rc = fetch(s.c01, C, 4, 1)) || (rc = fetch(s.c10, C, 4, 2)) ||
              (rc = fetch(s.c11, C, 4, 3)))) {
        return rc;
    }
    if (F && ((rc = fetch(s.f00, F, 4, 0)) || (rc = fetch(s.f01, F, 4, 1)) || (rc = fetch(s.f10, F, 4, 2)) ||
              (rc = fetch(s.f11, F, 4, 3)))) {
        return rc;
    }
    if (J && (rc = fetch(s.J, J, 1, 0))) { return rc; }
    if (material) {
        std::vector<uint8_t> m(N);
        HIP_TRY(h, hipMemcpy(m.data(), s.mat, N, hipMemcpyDeviceToHost));
        for (size_t j = 0; j < N; ++j) { material[ids[j]] = m[j]; }
    }
    return MPM_OK;
}

int32_t mpm_save_state(MpmEngine* h, const char* path)
{
    if (!h || !path) { return MPM_ERR_INVALID_ARG; }
    int32_t cnt = 0;
    int32_t rc = mpm_count(h, &cnt);
    if (rc != MPM_OK) { return rc; }
    const size_t N = (size_t)cnt;
    StatePayload sp(N);
    if (N > 0) {
        rc = mpm_readback(h, sp.x, sp.v, sp.C, sp.F, sp.J, sp.mat, nullptr);
        if (rc != MPM_OK) { return rc; }
        const Store& st = h->store[h->cur];
        std::vector<int32_t> ids(N);
        std::vector<uint32_t> keys(N);
        HIP_TRY(h, hipMemcpy(ids.data(), st.id, N * sizeof(int32_t), hipMemcpyDeviceToHost));
        HIP_TRY(h, hipMemcpy(keys.data(), st.key, N * sizeof(uint32_t), hipMemcpyDeviceToHost));
        for (size_t j = 0; j < N; ++j) {
            const int32_t sc = h->params.num_scenes > 1 ? (int32_t)(keys[j] / (uint32_t)h->dp.bps) : 0;
            std::memcpy(sp.scene + 4 * (size_t)ids[j], &sc, 4);
        }
    }
    StateHeader hd{};
    std::memcpy(hd.magic, kStateMagic, 8);
    hd.version = kStateVersion;
    hd.header_bytes = (uint32_t)sizeof(StateHeader);
    hd.n = cnt;
    hd.num_scenes = h->params.num_scenes;
    hd.checksum = fnv1a64(sp.bytes.data(), sp.bytes.size());
    hd.params = h->params;
    FILE* f = std::fopen(path, "wb");
    if (!f) { return fail(h, MPM_ERR_INVALID_ARG, 0, std::string("cannot open ") + path + " for writing"); }
    bool ok = std::fwrite(&hd, sizeof(hd), 1, f) == 1;
    ok = ok && (sp.bytes.empty() || std::fwrite(sp.bytes.data(), 1, sp.bytes.size(), f) == sp.bytes.size());
    ok = (std::fclose(f) == 0) && ok;
    return ok ? MPM_OK : fail(h, MPM_ERR_INVALID_ARG, 0, std::string("write failed: ") + path);
}

int32_t mpm_load_state(MpmEngine* h, const char* path)
{
    if (!h || !path) { return MPM_ERR_INVALID_ARG; }
    FILE* f = std::fopen(path, "rb");
    if (!f) { return fail(h, MPM_ERR_INVALID_ARG, 0, std::string("cannot open ") + path); }
    StateHeader hd{};
    const bool got = std::fread(&hd, sizeof(hd), 1, f) == 1;
    auto bad = [&](const std::string& why) {
        std::fclose(f);
        return fail(h, MPM_ERR_INVALID_ARG, 0, std::string("state file ") + path + ": " + why);
    };
    if (!got || std::memcmp(hd.magic, kStateMagic, 8) != 0) { return bad("not an MPM state file"); }
    if (hd.version != kStateVersion || hd.header_bytes != sizeof(StateHeader)) { return bad("unsupported version"); }
    const MpmParams& p = h->params;
    if (hd.num_scenes != p.num_scenes || hd.params.grid_w != p.grid_w || hd.params.grid_h != p.grid_h) {
        return bad("scene count or grid size differs from the handle's");
    }
    if (hd.n < 0) { return bad("negative particle count"); }
    if (hd.n > p.capacity) {
        std::fclose(f);
        return fail(h, MPM_ERR_CAPACITY, 0, "state file holds more particles than the handle's capacity");
    }
    StatePayload sp((size_t)hd.n);
    if (!sp.bytes.empty() && std::fread(sp.bytes.data(), 1, sp.bytes.size(), f) != sp.bytes.size()) {
        return bad("truncated");
    }
    if (std::fgetc(f) != EOF) { return bad("trailing bytes"); }
    std::fclose(f);
    if (fnv1a64(sp.bytes.data(), sp.bytes.size()) != hd.checksum) {
        return fail(h, MPM_ERR_INVALID_ARG, 0, std::string("state file ") + path + ": checksum mismatch");
    }
    std::vector<uint32_t> scene((size_t)hd.n);
    for (size_t i = 0; i < (size_t)hd.n; ++i) {
        int32_t sc;
        std::memcpy(&sc, sp.scene + 4 * i, 4);
        if (sc < 0 || sc >= p.num_scenes) { return fail(h, MPM_ERR_INVALID_ARG, 0, "state file: scene out of range"); }
        scene[i] = (uint32_t)sc;
    }
    // float arrays live at 4-byte aligned offsets of the payload
    return set_particles_impl(h, hd.n, sp.x, sp.v, sp.C, sp.F, sp.J, sp.mat, scene.data());
}

int32_t mpm_read_grid(MpmEngine* h, float* mass, float* vel_x, float* vel_y)
{
    if (!h) { return MPM_ERR_INVALID_ARG; }
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    const size_t nodes = (size_t)h->dp.G * (size_t)h->params.num_scenes;
    if (mass) { HIP_TRY(h, hipMemcpy(mass, h->grid.vmass, nodes * sizeof(float), hipMemcpyDeviceToHost)); }
    if (vel_x || vel_y) {
        std::vector<float2> v(nodes);
        HIP_TRY(h, hipMemcpy(v.data(), h->grid.vel, nodes * sizeof(float2), hipMemcpyDeviceToHost));
        for (size_t k = 0; k < nodes; ++k) {
            if (vel_x) { vel_x[k] = v[k].x; }
            if (vel_y) { vel_y[k] = v[k].y; }
        }
    }
    return MPM_OK;
}

int32_t mpm_hose_pipe(MpmEngine* h, float origin_x, float origin_y, float velocity_x, float velocity_y,
                      float nozzle_width, float spacing, float jitter, uint64_t tick, int32_t* emitted)
{
    if (!h) { return MPM_ERR_INVALID_ARG; }
    if (!(spacing > 0.f) || !(nozzle_width >= 0.f) || !(jitter >= 0.f)) {
        return fail(h, MPM_ERR_INVALID_ARG, 0, "spacing must be > 0, width and jitter >= 0");
    }
    // hosePipe, MPM.cu:450-453
    const float speed = std::sqrt(velocity_x * velocity_x + velocity_y * velocity_y);
    const int32_t width = std::max(1, (int32_t)std::lround(nozzle_width / spacing));
    const int32_t rows = std::max(1, (int32_t)std::lround(speed / spacing));
    const int32_t count = width * rows;
    const float tick_len = (float)h->params.substeps * h->params.dt;
    const hipError_t e = launch_emitter(h->stream, h->dp, h->store[h->cur], h->count, h->params.capacity, origin_x,
                                        origin_y, velocity_x, velocity_y, width, spacing, jitter, count, tick,
                                        1.f / tick_len);
    if (e != hipSuccess) { return fail(h, MPM_ERR_DEVICE, (int32_t)e, "emitter launch"); }
    // the device keeps the first capacity - count particles (k_emitter), so the host count stays exact
    h->last_count = (int32_t)std::min<int64_t>(h->params.capacity, (int64_t)h->last_count + count);
    h->need_sort = true;  // the new particles are in no bin yet
    if (emitted) { *emitted = count; }
    return MPM_OK;
}

int32_t mpm_set_closure(MpmEngine* h, const float* weights)
{
    if (!h || !weights) { return MPM_ERR_INVALID_ARG; }
    for (int i = 0; i < 4675; ++i) {
        if (!std::isfinite(weights[i])) { return fail(h, MPM_ERR_INVALID_ARG, 0, "closure weights not finite"); }
    }
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    HIP_TRY(h, hipMemcpy(h->closure_w, weights, 4675 * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(h, launch_pack_closure(h->stream, h->closure_w, h->closure_w2));
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    h->closure_loaded = true;
    return MPM_OK;
}

int32_t mpm_autotune(const MpmParams* base, int32_t n, const float* x, const float* v, const uint8_t* material,
                     const int32_t* scene_counts, int32_t substeps, MpmTuneResult* out)
{
    if (!base || !out || substeps < 1 || n < 0 || (n > 0 && !x)) {
        return fail(nullptr, MPM_ERR_INVALID_ARG, 0, "mpm_autotune: bad arguments");
    }
    if (base->mode == MPM_MODE_REFERENCE) { return fail(nullptr, MPM_ERR_STATE, 0, "reference mode has no tiles"); }
    // Time one configuration: warm-up tick, then one tick measured with the handle's t0/t1 events.
    auto time_one = [&](const MpmParams& p, float* ms) -> int32_t {
        MpmEngine* h = nullptr;
        int32_t rc = mpm_create(&p, &h);
        if (rc != MPM_OK) { return rc; }
        rc = mpm_set_particles(h, n, x, v, nullptr, nullptr, nullptr, material, scene_counts);
        if (rc == MPM_OK) { rc = mpm_step(h, 1); }
        if (rc == MPM_OK) { rc = mpm_step(h, 1); }
        if (rc == MPM_OK) { rc = mpm_last_step_ms(h, ms); }
        if (rc != MPM_OK) {
            g_create_error = h->last_error;
            g_create_msg = h->last_msg;
        }
        mpm_destroy(h);
        return rc;
    };
    MpmParams best = *base;
    best.substeps = substeps;
    if (best.sort_interval <= 0 || best.sort_interval > substeps) { best.sort_interval = substeps; }
    float best_ms = 3.4e38f;
    int32_t tried = 0;
    // 8192: one work item for bins of up to 8192 particles (16 x 16 bins at 16 particles per cell)
    const int chunks[5] = {512, 1024, 2048, 4096, 8192};
    for (int k = 0; k < kNumShapes; ++k) {
        for (int c : chunks) {
            MpmParams p = best;
            p.block_x = kShapeX[k];
            p.block_y = kShapeY[k];
            p.chunk = c;
            float ms = 0.f;
            const int32_t rc = time_one(p, &ms);
            if (rc != MPM_OK) { return rc; }
            ++tried;
            if (ms < best_ms) {
                best_ms = ms;
                best.block_x = p.block_x;
                best.block_y = p.block_y;
                best.chunk = c;
            }
        }
    }
    const int intervals[3] = {25, 50, 100};
    for (int iv : intervals) {
        if (iv > substeps || iv == best.sort_interval) { continue; }
        MpmParams p = best;
        p.sort_interval = iv;
        float ms = 0.f;
        const int32_t rc = time_one(p, &ms);
        if (rc != MPM_OK) { return rc; }
        ++tried;
        if (ms < best_ms) {
            best_ms = ms;
            best.sort_interval = iv;
        }
    }
    out->block_x = best.block_x;
    out->block_y = best.block_y;
    out->chunk = best.chunk;
    out->sort_interval = best.sort_interval;
    out->ms_per_substep = best_ms / (float)substeps;
    out->candidates = tried;
    return MPM_OK;
}

int32_t mpm_last_error(MpmEngine* h) { return h ? h->last_error : g_create_error; }

const char* mpm_error_name(int32_t err) { return hipGetErrorName((hipError_t)err); }

const char* mpm_status_name(int32_t status)
{
    switch (status) {
    case MPM_OK: return "MPM_OK";
    case MPM_BUSY: return "MPM_BUSY";
    case MPM_ERR_DEVICE: return "MPM_ERR_DEVICE";
    case MPM_ERR_NO_DEVICE: return "MPM_ERR_NO_DEVICE";
    case MPM_ERR_STATE: return "MPM_ERR_STATE";
    case MPM_ERR_INVALID_ARG: return "MPM_ERR_INVALID_ARG";
    case MPM_ERR_ALREADY: return "MPM_ERR_ALREADY";
    case MPM_ERR_CAPACITY: return "MPM_ERR_CAPACITY";
    case MPM_ERR_NOT_BUILT: return "MPM_ERR_NOT_BUILT";
    default: return "MPM_UNKNOWN_STATUS";
    }
}

const char* mpm_last_message(MpmEngine* h) { return h ? h->last_msg.c_str() : g_create_msg.c_str(); }

int32_t mpm_debug_bin(MpmEngine* h, uint32_t* keys, int32_t* ids, int32_t* n)
{
    if (!h) { return MPM_ERR_INVALID_ARG; }
    if (h->params.mode == MPM_MODE_REFERENCE) { return fail(h, MPM_ERR_STATE, 0, "no binning in reference mode"); }
    int32_t cnt = 0;
    int32_t rc = mpm_count(h, &cnt);
    if (rc != MPM_OK) { return rc; }
    const int c = h->cur;
    HIP_TRY(h, launch_binning(h->stream, h->dp, h->store[c], h->store[c ^ 1], h->count, h->ss));
    h->cur ^= 1;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    const Store& s = h->store[h->cur];
    if (keys && cnt) { HIP_TRY(h, hipMemcpy(keys, s.key, (size_t)cnt * sizeof(uint32_t), hipMemcpyDeviceToHost)); }
    if (ids && cnt) { HIP_TRY(h, hipMemcpy(ids, s.id, (size_t)cnt * sizeof(int32_t), hipMemcpyDeviceToHost)); }
    if (n) { *n = cnt; }
    return MPM_OK;
}

int32_t mpm_time_kernels(MpmEngine* h, int32_t substeps, float* ms, int32_t* launches)
{
    if (!h || !ms || !launches || substeps < 1) { return MPM_ERR_INVALID_ARG; }
    if (h->params.closure == MPM_CLOSURE_LEARNED && !h->closure_loaded) {
        return fail(h, MPM_ERR_STATE, 0, "learned closure selected but no weights loaded (mpm_set_closure)");
    }
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    LaunchTimer tm;
    tm.st = h->stream;
    int c = h->cur;
    bool dirty[3] = {false, false, false};
    // whole ticks follow the binning schedule of mpm_step; a remainder shorter than a tick bins first
    hipError_t e = hipSuccess;
    for (int left = substeps; e == hipSuccess && left > 0;) {
        const int n = std::min(left, h->params.substeps);
        BinStart bin = next_tick_bins(h);
        if (n < h->params.substeps && bin == kBinNone) { bin = kBinFirst; }
        e = enqueue_substeps(h, c, n, false, &tm, dirty, bin);
        note_tick(h, bin != kBinNone);
        left -= n;
    }
    if (e == hipSuccess) { e = tm.close(); }
    if (e != hipSuccess) { return fail(h, MPM_ERR_DEVICE, (int32_t)e, "timed launch"); }
    h->cur = c;
    HIP_TRY(h, hipStreamSynchronize(h->stream));
    for (int k = 0; k < MPM_KT_COUNT; ++k) {
        ms[k] = 0.f;
        launches[k] = 0;
    }
    for (size_t i = 0; i < tm.cls.size(); ++i) {
        float t = 0.f;
        HIP_TRY(h, hipEventElapsedTime(&t, tm.ev[2 * i], tm.ev[2 * i + 1]));
        ms[tm.cls[i]] += t;
        launches[tm.cls[i]] += tm.count[i];
    }
    return MPM_OK;
}

namespace {
// Stress diagnostics on host arrays (row-major F[4n], J[n], tau[3n]); snow_return: snow particles take the
// return mapping first and F_out / J_out (host, may be NULL without snow_return) receive the result.
int32_t stress_impl(MpmEngine* h, int32_t n, const uint8_t* material, const float* F, const float* J, float* tau,
                    int snow_return, float* F_out, float* J_out)
{
    const size_t N = (size_t)n;
    std::vector<float> planes(5 * N);
    for (size_t i = 0; i < N; ++i) {
        for (int k = 0; k < 4; ++k) { planes[k * N + i] = F[4 * i + k]; }
        planes[4 * N + i] = J[i];
    }
    uint8_t* buf = nullptr;
    const size_t bytes = 13 * N * sizeof(float) + N;
    HIP_TRY(h, hipMalloc((void**)&buf, bytes));
    float* dF = reinterpret_cast<float*>(buf);
    float* dJ = dF + 4 * N;
    float* dT = dJ + N;
    float* dFo = dT + 3 * N;
    float* dJo = dFo + 4 * N;
    uint8_t* dM = reinterpret_cast<uint8_t*>(dJo + N);
    hipError_t e = hipMemcpy(dF, planes.data(), 5 * N * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) { e = hipMemcpy(dM, material, N, hipMemcpyHostToDevice); }
    if (e == hipSuccess) { e = launch_stress(h->stream, h->dp, n, dM, dF, dJ, dT, snow_return, dFo, dJo); }
    if (e == hipSuccess) { e = hipStreamSynchronize(h->stream); }
    std::vector<float> t(8 * N);
    if (e == hipSuccess) { e = hipMemcpy(t.data(), dT, 8 * N * sizeof(float), hipMemcpyDeviceToHost); }
    (void)hipFree(buf);
    if (e != hipSuccess) { return fail(h, MPM_ERR_DEVICE, (int32_t)e, "stress kernel"); }
    for (size_t i = 0; i < N; ++i) {
        for (int k = 0; k < 3; ++k) { tau[3 * i + k] = t[k * N + i]; }
        if (F_out) {
            for (int k = 0; k < 4; ++k) { F_out[4 * i + k] = t[(3 + k) * N + i]; }
        }
        if (J_out) { J_out[i] = t[7 * N + i]; }
    }
    return MPM_OK;
}
}  // namespace

int32_t mpm_compute_stress(MpmEngine* h, int32_t n, const uint8_t* material, const float* F, const float* J,
                           float* tau)
{
    if (!h || n < 0 || (n > 0 && (!material || !F || !J || !tau))) { return MPM_ERR_INVALID_ARG; }
    if (n == 0) { return MPM_OK; }
    return stress_impl(h, n, material, F, J, tau, 0, nullptr, nullptr);
}

int32_t mpm_compute_snow_return(MpmEngine* h, int32_t n, const float* F_trial, const float* Jp, float* F_out,
                                float* Jp_out, float* tau)
{
    if (!h || n < 0 || (n > 0 && (!F_trial || !Jp || !F_out || !Jp_out || !tau))) { return MPM_ERR_INVALID_ARG; }
    if (n == 0) { return MPM_OK; }
    std::vector<uint8_t> snow((size_t)n, (uint8_t)MPM_MAT_SNOW);
    return stress_impl(h, n, snow.data(), F_trial, Jp, tau, 1, F_out, Jp_out);
}

}  // extern "C"

