// facade_shim.cpp -- test driver of the C++20 host side (include/mitrace_gpu_tracer.hpp).
//
// Each entry point below runs the facade the way the reference's callers run its Tracer
// (RenderThread: construct, StartRender, WaitForRender; StatThread: GetStats polled from another
// thread while the render runs; the kill function from a third party, src/main.cpp:41-142) and
// reports what it saw through plain C arguments, so the Python test-suite can compare the
// result with the C-ABI's synchronous mt_render and with the oracle.  Errors thrown by the
// facade are caught here and returned as their MT_E_* code with the message.
#include <atomic>
#include <chrono>
#include <cstring>
#include <thread>

#include "../../include/mitrace_gpu_tracer.hpp"

namespace {

void copy_message(const char* msg, char* err, int errlen) {
    if (!err || errlen <= 0) return;
    std::strncpy(err, msg, static_cast<std::size_t>(errlen) - 1);
    err[errlen - 1] = '\0';
}

}  // namespace

extern "C" {

int fac_default_settings(mt_settings* out) {
    *out = mitrace_gpu::DefaultSettings();
    return 0;
}

// Construct + StartRender + WaitForRender with a StatThread-like poller.  stats_out: rays,
// samples after the wait; polls_out: [0] polls made, [1] polls whose progress was in [0, 1] and
// not below the previous one, [2] 1 if IsDone() was false before the start and true after the
// wait.
int fac_render(const mt_scene_desc* desc, const mt_camera_desc* camera, const mt_settings* settings,
               float* rgb_out, uint64_t* stats_out, int* polls_out, char* err, int errlen) {
    try {
        mitrace_gpu::Scene scene(*desc, *camera);
        auto buffer = std::make_shared<mitrace_gpu::RenderBuffer>(settings->width, settings->height);
        mitrace_gpu::Tracer tracer(buffer, scene, *settings);
        const bool done_before = tracer.IsDone();
        tracer.StartRender();
        std::atomic<bool> stop{false};
        int polls = 0, monotone = 0;
        std::thread stat([&] {
            float last = 0.0f;
            while (!stop.load()) {
                const auto s = tracer.GetStats();
                ++polls;
                if (s.progress >= last && s.progress >= 0.0f && s.progress <= 1.0f) ++monotone;
                last = s.progress;
                std::this_thread::sleep_for(std::chrono::milliseconds(1));
            }
        });
        tracer.WaitForRender();
        stop.store(true);
        stat.join();
        const auto s = tracer.GetStats();
        stats_out[0] = s.rays;
        stats_out[1] = s.samples;
        polls_out[0] = polls;
        polls_out[1] = monotone;
        polls_out[2] = (!done_before && tracer.IsDone()) ? 1 : 0;
        std::memcpy(rgb_out, buffer->Data(),
                    sizeof(float) * 3 * std::size_t(settings->width) * settings->height);
        return 0;
    } catch (const mitrace_gpu::Error& e) {
        copy_message(e.what(), err, errlen);
        return e.Code();
    }
}

// StartRender, KillRender from another thread once the first chunk has finished, then
// WaitForRender: returns 0 when the facade treats the kill as a normal end.  out: [0] samples
// rendered, [1] IsDone() after the wait, [2] 1 if a second StartRender after the killed one
// completed the full sample count.
int fac_kill(const mt_scene_desc* desc, const mt_camera_desc* camera, const mt_settings* settings,
             uint64_t* out, char* err, int errlen) {
    try {
        mitrace_gpu::Scene scene(*desc, *camera);
        auto buffer = std::make_shared<mitrace_gpu::RenderBuffer>(settings->width, settings->height);
        mitrace_gpu::Tracer tracer(buffer, scene, *settings);
        tracer.StartRender();
        std::thread killer([&] {
            for (int i = 0; i < 20000 && tracer.GetStats().progress <= 0.0f && !tracer.IsDone(); ++i)
                std::this_thread::sleep_for(std::chrono::microseconds(200));
            tracer.KillRender();
        });
        tracer.WaitForRender();
        killer.join();
        out[0] = tracer.GetStats().samples;
        out[1] = tracer.IsDone() ? 1 : 0;
        tracer.StartRender();  // the killed job is released and a new render runs to the end
        tracer.WaitForRender();
        const uint64_t full = uint64_t(settings->width) * settings->height * settings->samples;
        out[2] = tracer.GetStats().samples == full ? 1 : 0;
        return 0;
    } catch (const mitrace_gpu::Error& e) {
        copy_message(e.what(), err, errlen);
        return e.Code();
    }
}

// Scene construction only: returns 0, or the code of the Error it threw (message in err).
int fac_scene_error(const mt_scene_desc* desc, char* err, int errlen) {
    try {
        mt_camera_desc cam{};
        mitrace_gpu::Scene scene(*desc, cam);
        return 0;
    } catch (const mitrace_gpu::Error& e) {
        copy_message(e.what(), err, errlen);
        return e.Code();
    }
}

// RenderBuffer::GetPixelsRGB8 of a W x H float RGB image.
int fac_rgb8(const float* pixels, unsigned w, unsigned h, uint8_t* out) {
    mitrace_gpu::RenderBuffer b(w, h);
    std::memcpy(b.Data(), pixels, sizeof(float) * 3 * std::size_t(w) * h);
    auto p = b.GetPixelsRGB8();
    std::memcpy(out, p.get(), 3 * std::size_t(w) * h);
    return 0;
}

// RenderBuffer::DrawSubBuffer: dst (dw x dh) receives src (sw x sh) at (ox, oy); result in out.
int fac_draw_sub(const float* dst, unsigned dw, unsigned dh, const float* src, unsigned sw,
                 unsigned sh, unsigned ox, unsigned oy, float* out) {
    mitrace_gpu::RenderBuffer a(dw, dh), b(sw, sh);
    std::memcpy(a.Data(), dst, sizeof(float) * 3 * std::size_t(dw) * dh);
    std::memcpy(b.Data(), src, sizeof(float) * 3 * std::size_t(sw) * sh);
    a.DrawSubBuffer(b, ox, oy);
    std::memcpy(out, a.Data(), sizeof(float) * 3 * std::size_t(dw) * dh);
    return 0;
}

// RenderBuffer::SaveToFilePNG (which = 0), SaveToFileHDR (which = 1) or SaveToFileJPG
// (which = 2) of a W x H float RGB image; returns 0, or 1 with the exception's message in err.
int fac_save(const float* pixels, unsigned w, unsigned h, const char* path, int which, char* err,
             int errlen) {
    try {
        mitrace_gpu::RenderBuffer b(w, h);
        std::memcpy(b.Data(), pixels, sizeof(float) * 3 * std::size_t(w) * h);
        if (which == 0)
            b.SaveToFilePNG(path);
        else if (which == 1)
            b.SaveToFileHDR(path);
        else
            b.SaveToFileJPG(path);
        return 0;
    } catch (const std::runtime_error& e) {
        copy_message(e.what(), err, errlen);
        return 1;
    }
}

}  // extern "C"
