// mpm_run — headless driver of the MPM engine over its C-ABI (include/mpm_engine.h, include/mpm_render.h).
//
// Counterpart of the reference's tick loop (main/MPM-master.cpp:33-125) without the window and GUI thread:
//   - hose-pipe emission every tick until TARGET_COUNT particles (MPM-master.cpp:19-25, 103-105),
//   - one engine tick per loop iteration, timed on the device (t0 / t1 events, MPM-master.cpp:106-109, 90-92),
//   - every 150 ticks the average and longest tick time and the particle count (MPM-master.cpp:94-102),
//   - optionally the frame of the reference renderer written as PPM images (cuda-render.cu:163-310).
// The reference paces ticks at 32 ms of wall time (kTickPeriod, constants.h:24); --realtime does the same,
// the default runs ticks back to back.
//
//   mpm_run [--ticks N] [--preset P] [--state FILE] [--save FILE] [--target N] [--frames-every K]
//           [--out DIR] [--realtime]
//
// Presets other than the reference one start from a state file (--state, written by mpm_save_state) and
// do not emit.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/mpm_engine.h"
#include "../../include/mpm_render.h"

namespace {

// hose-pipe settings of MPM-master.cpp:19-25
constexpr float kOriginX = 30.f, kOriginY = 30.f;
constexpr float kVelocityX = 12.f, kVelocityY = 0.f;
constexpr float kNozzleWidth = 30.f;
constexpr float kSpacing = 1.05f * 1.2f;
constexpr float kJitter = kSpacing * 0.2f;
constexpr int kSampleTicks = 150;

struct Options {
    int ticks = 600;
    int preset = MPM_PRESET_REFERENCE;
    int target = 100000;
    int frames_every = 0;
    bool realtime = false;
    std::string state, save, out = ".";
};

[[noreturn]] void usage(const char* why)
{
    std::fprintf(stderr,
                 "mpm_run: %s\nusage: mpm_run [--ticks N] [--preset P] [--state FILE] [--save FILE] [--target N]\n"
                 "               [--frames-every K] [--out DIR] [--realtime]\n",
                 why);
    std::exit(2);
}

Options parse(int argc, char** argv)
{
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> const char* {
            if (i + 1 >= argc) { usage(("missing value after " + a).c_str()); }
            return argv[++i];
        };
        if (a == "--ticks") {
            o.ticks = std::atoi(next());
        } else if (a == "--preset") {
            o.preset = std::atoi(next());
        } else if (a == "--target") {
            o.target = std::atoi(next());
        } else if (a == "--frames-every") {
            o.frames_every = std::atoi(next());
        } else if (a == "--state") {
            o.state = next();
        } else if (a == "--save") {
            o.save = next();
        } else if (a == "--out") {
            o.out = next();
        } else if (a == "--realtime") {
            o.realtime = true;
        } else {
            usage(("unknown option " + a).c_str());
        }
    }
    if (o.ticks < 0 || o.frames_every < 0 || o.target < 0) { usage("negative count"); }
    return o;
}

void check(MpmEngine* h, int32_t rc, const char* what)
{
    if (rc < 0) {
        std::fprintf(stderr, "mpm_run: %s failed: %s (%d): %s\n", what, mpm_status_name(rc), rc, mpm_last_message(h));
        std::exit(1);
    }
}

void check_render(int32_t rc, const char* what)
{
    if (rc < 0) {
        std::fprintf(stderr, "mpm_run: %s failed: %s (%d): %s\n", what, mpm_status_name(rc), rc, mpm_render_message());
        std::exit(1);
    }
}

}  // namespace

int main(int argc, char** argv)
{
    const Options o = parse(argc, argv);
    MpmParams p;
    if (mpm_params_preset(o.preset, &p) != MPM_OK) { usage("unknown preset"); }
    const bool emit = o.preset == MPM_PRESET_REFERENCE && o.state.empty();
    if (!emit && o.state.empty()) { usage("presets other than 0 need --state FILE"); }
    MpmEngine* h = nullptr;
    check(nullptr, mpm_create(&p, &h), "mpm_create");
    if (!o.state.empty()) { check(h, mpm_load_state(h, o.state.c_str()), "mpm_load_state"); }

    MpmRenderer* r = nullptr;
    MpmRenderDesc rd{};
    std::vector<uint8_t> image;
    if (o.frames_every > 0) {
        if (o.preset == MPM_PRESET_REFERENCE) {
            // reference display: kDispWidth x kDispHeight, kSimScale = 5, chunk lines every 32 * 5 px (constants.h:48-61)
            rd = MpmRenderDesc{1920, 960, 0, 160, 5.f, 0.f};
        } else {
            // unit-square scenes with y up: one pixel per grid cell, flipped
            const int w = p.grid_w - 2 * p.halo, hh = p.grid_h - 2 * p.halo;
            rd = MpmRenderDesc{w, hh, 0, 0, (float)w / p.sim_w, p.sim_h};
        }
        check_render(mpm_renderer_create(&rd, &r), "mpm_renderer_create");
        image.resize((size_t)4 * rd.width * rd.height);
    }

    int32_t atoms = 0;
    check(h, mpm_count(h, &atoms), "mpm_count");
    float cum = 0.f, longest = 0.f, all_ms = 0.f, all_longest = 0.f;
    int timed = 0, frames = 0;
    uint64_t last_sampled = 0;
    auto next_tick = std::chrono::steady_clock::now();
    for (uint64_t tick = 0; tick <= (uint64_t)o.ticks; ++tick) {
        if (tick > 0) {
            check(h, mpm_sync(h), "mpm_sync");
            float ms = 0.f;
            check(h, mpm_last_step_ms(h, &ms), "mpm_last_step_ms");
            cum += ms;
            longest = ms > longest ? ms : longest;
            all_ms += ms;
            all_longest = ms > all_longest ? ms : all_longest;
            ++timed;
            if (r && tick % (uint64_t)o.frames_every == 0) {
                MpmView view;
                check(h, mpm_read_view(h, &view), "mpm_read_view");
                check_render(mpm_render_frame(r, &view, image.data(), 0), "mpm_render_frame");
                char name[64];
                std::snprintf(name, sizeof(name), "/frame_%05u.ppm", (unsigned)tick);
                check_render(mpm_write_ppm((o.out + name).c_str(), image.data(), rd.width, rd.height, 0), "mpm_write_ppm");
                ++frames;
            }
        }
        if (tick - last_sampled == (uint64_t)kSampleTicks) {
            std::printf("------- Tick %u -------\nAverage duration = %f ms\nLongest tick = %f ms\nSimulating %d atoms\n",
                        (unsigned)tick, cum / (float)kSampleTicks, longest, atoms);
            std::fflush(stdout);
            cum = 0.f;
            longest = 0.f;
            last_sampled = tick;
        }
        if (tick == (uint64_t)o.ticks) { break; }
        if (emit && atoms < o.target) {
            int32_t emitted = 0;
            check(h, mpm_hose_pipe(h, kOriginX, kOriginY, kVelocityX, kVelocityY, kNozzleWidth, kSpacing, kJitter, tick,
                                   &emitted),
                  "mpm_hose_pipe");
            atoms += emitted;
        }
        check(h, mpm_step(h, 1), "mpm_step");
        if (o.realtime) {
            next_tick += std::chrono::milliseconds(32);
            std::this_thread::sleep_until(next_tick);
        }
    }
    check(h, mpm_sync(h), "mpm_sync");
    check(h, mpm_count(h, &atoms), "mpm_count");
    if (!o.save.empty()) { check(h, mpm_save_state(h, o.save.c_str()), "mpm_save_state"); }
    std::printf("mpm_run: ticks=%d particles=%d mean_tick_ms=%.4f longest_tick_ms=%.4f frames=%d\n", timed, atoms,
                timed ? all_ms / (float)timed : 0.f, all_longest, frames);
    if (r) { mpm_renderer_destroy(r); }
    mpm_destroy(h);
    return 0;
}
