/*
 * mitrace_gpu_tracer.hpp -- C++20 host side over the C-ABI of mitrace_gpu.h, with the class
 * and method set of the reference's CPU Tracer, so that its callers (RenderThread and
 * StatThread, src/main.cpp:41-142) change only the type they name.  Header-only; link
 * libmitrace_gpu.so.
 *
 *   this header                     reference
 *   ------------------------------  ---------------------------------------------------------
 *   mitrace_gpu::RenderBuffer       RenderBuffer (src/Tracer/RenderBuffer.hpp:11-35,
 *                                     RenderBuffer.cpp:3-84): W x H float RGB, row-major;
 *                                     SaveToFilePNG / SaveToFileHDR
 *   mitrace_gpu::Scene              Scene::Scene (src/Scene/Scene.cpp:53-59) as uploaded by
 *                                     mt_scene_create; GetCamera() as Scene::GetCamera
 *   mitrace_gpu::DefaultSettings()  Config::ConfigOptions defaults (src/CLI/Config.hpp:43-80)
 *   mitrace_gpu::Tracer             Tracer::Tracer (src/Tracer/Tracer.hpp:19-129):
 *     Tracer(buffer, scene)           Tracer.cpp:633-640 (settings passed in, not read from
 *                                     the Config singleton)
 *     StartRender()                   Tracer.cpp:718-779  -> mt_render_start
 *     WaitForRender()                 Tracer.cpp:781-799  -> mt_render_wait
 *     KillRender()                    Tracer.cpp:801-804  -> mt_render_kill
 *     GetStats()                      Tracer.cpp:806-821  -> mt_render_stats
 *     IsDone()                        Tracer.cpp:823-828  -> mt_render_is_done
 *
 * Errors: a failing C call throws mitrace_gpu::Error, a std::runtime_error that carries the
 * MT_E_* code and the message of mt_last_error(), which keeps the reference's convention of
 * runtime_error for load and render failures (src/main.cpp:29-30).  A killed render is not an
 * error: WaitForRender returns, as the reference's does after KillRender.
 *
 * Threading, as in the reference: StartRender does not block; GetStats, IsDone and KillRender
 * may be called from other threads while a render runs (StatThread polls GetStats every
 * second); construction, StartRender, WaitForRender and destruction belong to one thread.
 */
#ifndef MITRACE_GPU_TRACER_HPP
#define MITRACE_GPU_TRACER_HPP

#include <algorithm>
#include <array>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mitrace_gpu.h"

namespace mitrace_gpu {

class Error : public std::runtime_error {
  public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int Code() const noexcept { return code_; }

  private:
    int code_;
};

/* Throws Error when a C-ABI call returned a negative MT_E_* code. */
inline void Check(int rc, const char* call) {
    if (rc == MT_OK) return;
    const char* msg = mt_last_error();
    std::string text = std::string(call) + " failed (" + std::to_string(rc) + ")";
    if (msg && *msg) text += ": " + std::string(msg);
    throw Error(rc, text);
}

using Color = std::array<float, 3>;

/* RenderBuffer (src/Tracer/RenderBuffer.hpp:11-35).  Data() is the one accessor the GPU path
 * adds: the library writes the finished W*H*3 floats there. */
class RenderBuffer {
  public:
    RenderBuffer(unsigned width, unsigned height)
        : width_(width), height_(height), pixels_(std::size_t(width) * height * 3, 0.0f) {}

    unsigned GetWidth() const noexcept { return width_; }
    unsigned GetHeight() const noexcept { return height_; }

    /* out-of-range coordinates are ignored, as RenderBuffer.cpp:3-24 does */
    void GetPixel(unsigned x, unsigned y, Color& color) const noexcept {
        if (x >= width_ || y >= height_) return;
        const float* p = &pixels_[(std::size_t(y) * width_ + x) * 3];
        color = {p[0], p[1], p[2]};
    }
    void SetPixel(unsigned x, unsigned y, const Color& color) noexcept {
        if (x >= width_ || y >= height_) return;
        float* p = &pixels_[(std::size_t(y) * width_ + x) * 3];
        p[0] = color[0], p[1] = color[1], p[2] = color[2];
    }
    void AddColorAt(unsigned x, unsigned y, const Color& color) noexcept {
        if (x >= width_ || y >= height_) return;
        float* p = &pixels_[(std::size_t(y) * width_ + x) * 3];
        p[0] += color[0], p[1] += color[1], p[2] += color[2];
    }
    /* RenderBuffer.cpp:26-37: copies the sub-buffer, clipping at this buffer's edges */
    void DrawSubBuffer(const RenderBuffer& sub, unsigned offsetX, unsigned offsetY) noexcept {
        for (unsigned y = 0; y < sub.GetHeight(); ++y)
            for (unsigned x = 0; x < sub.GetWidth(); ++x) {
                Color c{};
                sub.GetPixel(x, y, c);
                SetPixel(offsetX + x, offsetY + y, c);
            }
    }
    /* RenderBuffer.cpp:39-58: trunc(clamp(v * 255, 0, 255)), no sRGB curve */
    std::shared_ptr<std::uint8_t[]> GetPixelsRGB8() const {
        std::shared_ptr<std::uint8_t[]> out(new std::uint8_t[pixels_.size()]);
        for (std::size_t i = 0; i < pixels_.size(); ++i)
            out[i] = static_cast<std::uint8_t>(std::clamp(pixels_[i] * 255.0f, 0.0f, 255.0f));
        return out;
    }

    /* RenderBuffer.cpp:60-84 write through stb; these writers are self-contained.  PNG: 8-bit
     * RGB of GetPixelsRGB8, zlib stream of stored (uncompressed) deflate blocks, so any PNG
     * reader decodes the same pixels stb's compressed file holds.  HDR: Radiance RGBE with
     * flat (not run-length encoded) scanlines, top row first, linear floats as the buffer
     * holds them.  A failure throws std::runtime_error("Failed to save image to file"), the
     * reference's message.  SaveToFileJPG has no counterpart here (no JPEG encoder). */
    void SaveToFilePNG(const std::string& filename) const {
        const auto rgb = GetPixelsRGB8();
        std::vector<std::uint8_t> raw;  // filter byte 0 (None) + one row
        raw.reserve((std::size_t(width_) * 3 + 1) * height_);
        for (unsigned y = 0; y < height_; ++y) {
            raw.push_back(0);
            raw.insert(raw.end(), rgb.get() + std::size_t(y) * width_ * 3,
                       rgb.get() + std::size_t(y + 1) * width_ * 3);
        }
        std::vector<std::uint8_t> z = {0x78, 0x01};
        std::size_t pos = 0;
        do {
            const std::size_t n = std::min<std::size_t>(65535, raw.size() - pos);
            z.push_back(pos + n == raw.size() ? 1 : 0);  // BFINAL, BTYPE 00 (stored)
            z.push_back(static_cast<std::uint8_t>(n & 0xff));
            z.push_back(static_cast<std::uint8_t>(n >> 8));
            z.push_back(static_cast<std::uint8_t>(~n & 0xff));
            z.push_back(static_cast<std::uint8_t>((~n >> 8) & 0xff));
            z.insert(z.end(), raw.begin() + pos, raw.begin() + pos + n);
            pos += n;
        } while (pos < raw.size());
        std::uint32_t a = 1, b = 0;  // Adler-32 of the uncompressed data (RFC 1950)
        for (std::uint8_t c : raw) {
            a = (a + c) % 65521u;
            b = (b + a) % 65521u;
        }
        PutU32(z, (b << 16) | a);
        std::vector<std::uint8_t> ihdr;
        PutU32(ihdr, width_);
        PutU32(ihdr, height_);
        ihdr.insert(ihdr.end(), {8, 2, 0, 0, 0});  // 8-bit, truecolour, deflate, no filter, no interlace
        std::vector<std::uint8_t> file = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
        PngChunk(file, "IHDR", ihdr);
        PngChunk(file, "IDAT", z);
        PngChunk(file, "IEND", {});
        WriteFile(filename, file);
    }

    void SaveToFileHDR(const std::string& filename) const {
        const std::string head = "#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y " + std::to_string(height_) +
                                 " +X " + std::to_string(width_) + "\n";
        std::vector<std::uint8_t> file(head.begin(), head.end());
        file.reserve(file.size() + pixels_.size() / 3 * 4);
        for (std::size_t i = 0; i < pixels_.size(); i += 3) {
            // RGBE: the largest component's binary exponent is shared (Ward, Graphics Gems II)
            const float r = pixels_[i], g = pixels_[i + 1], bl = pixels_[i + 2];
            const float m = std::max(r, std::max(g, bl));
            if (!(m >= 1e-32f)) {
                file.insert(file.end(), {0, 0, 0, 0});
                continue;
            }
            int e = 0;
            const float scale = std::frexp(m, &e) * 256.0f / m;
            file.push_back(static_cast<std::uint8_t>(std::max(r, 0.0f) * scale));
            file.push_back(static_cast<std::uint8_t>(std::max(g, 0.0f) * scale));
            file.push_back(static_cast<std::uint8_t>(std::max(bl, 0.0f) * scale));
            file.push_back(static_cast<std::uint8_t>(e + 128));
        }
        WriteFile(filename, file);
    }

    float* Data() noexcept { return pixels_.data(); }
    const float* Data() const noexcept { return pixels_.data(); }

  private:
    static void PutU32(std::vector<std::uint8_t>& v, std::uint32_t x) {
        for (int s = 24; s >= 0; s -= 8) v.push_back(static_cast<std::uint8_t>(x >> s));
    }
    static void PngChunk(std::vector<std::uint8_t>& file, const char* type,
                         const std::vector<std::uint8_t>& data) {
        PutU32(file, static_cast<std::uint32_t>(data.size()));
        const std::size_t start = file.size();
        file.insert(file.end(), type, type + 4);
        file.insert(file.end(), data.begin(), data.end());
        std::uint32_t crc = 0xffffffffu;  // CRC-32 (ISO 3309) over type and data
        for (std::size_t i = start; i < file.size(); ++i) {
            crc ^= file[i];
            for (int k = 0; k < 8; ++k) crc = (crc >> 1) ^ (0xedb88320u & (0u - (crc & 1u)));
        }
        PutU32(file, ~crc);
    }
    static void WriteFile(const std::string& filename, const std::vector<std::uint8_t>& bytes) {
        std::FILE* f = std::fopen(filename.c_str(), "wb");
        bool ok = f && std::fwrite(bytes.data(), 1, bytes.size(), f) == bytes.size();
        if (f) ok = (std::fclose(f) == 0) && ok;
        if (!ok) throw std::runtime_error("Failed to save image to file");
    }

    unsigned width_;
    unsigned height_;
    std::vector<float> pixels_;
};

/* Config::ConfigOptions defaults (src/CLI/Config.hpp:43-80; the firefly flag is inverted on
 * the command line, Config.cpp:166-170, so its effective default is off), with the GPU
 * additions at their neutral values: reference estimator, per-sample tone mapping, the
 * reference traversal, automatic wavefront size. */
inline mt_settings DefaultSettings() {
    mt_settings s{};
    s.width = 1280;
    s.height = 720;
    s.samples = 64;
    s.bounces = 5;
    s.ev_exposure = 8.0f;
    s.hdri_rotation = 0.0f;
    s.hdri_secondary_intensity = 350.0f;
    s.hdri_primary_intensity = 350.0f;
    s.emission_base_intensity = 1.0f;
    s.firefly_enable = 0;
    s.firefly_threshold = 6.0f;
    s.terminate_energy = 0.01f;
    s.image_block_size = 32;
    s.seed = 0;
    s.mode = MT_MODE_PARITY;
    s.rr_depth = 3;
    s.output = MT_OUTPUT_TONEMAPPED;
    s.hdri_rotation_apply = 0;
    s.paths_in_flight = 0;
    s.count_work = 0;
    s.debug_mode = MT_DEBUG_NONE;
    s.traversal = MT_TRAVERSAL_REFERENCE;
    return s;
}

/* The scene in HBM (mt_scene_create: BVH build with the reference's rules, flattening,
 * upload).  The description's arrays are copied; the caller may free them afterwards.  Like
 * Scene::Scene, it must outlive every Tracer that renders it. */
class Scene {
  public:
    Scene(const mt_scene_desc& desc, const mt_camera_desc& camera) : camera_(camera) {
        Check(mt_scene_create(&desc, &handle_), "mt_scene_create");
    }
    ~Scene() {
        if (handle_) mt_scene_destroy(handle_);
    }
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;
    Scene(Scene&& o) noexcept : handle_(std::exchange(o.handle_, nullptr)), camera_(o.camera_) {}
    Scene& operator=(Scene&& o) noexcept {
        if (this != &o) {
            if (handle_) mt_scene_destroy(handle_);
            handle_ = std::exchange(o.handle_, nullptr);
            camera_ = o.camera_;
        }
        return *this;
    }

    const mt_camera_desc& GetCamera() const noexcept { return camera_; }
    mt_scene* Handle() const noexcept { return handle_; }

  private:
    mt_scene* handle_ = nullptr;
    mt_camera_desc camera_;
};

/* Tracer::Tracer (src/Tracer/Tracer.hpp:19-129) on the GPU.  The camera is copied from the
 * scene at construction (Tracer.cpp:634), the settings likewise. */
class Tracer {
  public:
    /* Tracer::Stats (Tracer.hpp:78-85) */
    struct Stats {
        float progress;
        float timeElapsed;
        float estimatedTimeRemaining;
        std::size_t rays;
        std::size_t samples;
    };

    Tracer(std::shared_ptr<RenderBuffer> imageBuffer, const Scene& scene,
           const mt_settings& settings = DefaultSettings())
        : imageBuffer_(std::move(imageBuffer)), scene_(scene), cam_(scene.GetCamera()),
          settings_(settings) {
        if (!imageBuffer_) throw Error(MT_E_INVALID, "Tracer: null image buffer");
        if (imageBuffer_->GetWidth() != settings_.width ||
            imageBuffer_->GetHeight() != settings_.height)
            throw Error(MT_E_INVALID, "Tracer: image buffer and settings sizes differ");
    }
    ~Tracer() { mt_job_destroy(job_); }  // stops and joins a render still running
    Tracer(const Tracer&) = delete;
    Tracer& operator=(const Tracer&) = delete;

    /* Non-blocking.  A finished (or killed) earlier render's job is released first. */
    void StartRender() {
        if (job_) {
            if (!mt_render_is_done(job_))
                throw Error(MT_E_INVALID, "Tracer::StartRender: a render is still running");
            mt_job_destroy(std::exchange(job_, nullptr));
        }
        mt_job* job = nullptr;
        Check(mt_render_start(scene_.Handle(), &cam_, &settings_, imageBuffer_->Data(), &job),
              "mt_render_start");
        job_ = job;
    }

    /* Blocks until the render ends; the image (debug views post-processed, as
     * ApplyDebugPostProcessing does at Tracer.cpp:797) is in the buffer afterwards.  A render
     * stopped by KillRender returns normally. */
    void WaitForRender() {
        if (!job_) return;
        int rc = mt_render_wait(job_);
        if (rc != MT_OK && rc != MT_E_KILLED) Check(rc, "mt_render_wait");
    }

    void KillRender() noexcept {
        if (job_) mt_render_kill(job_);
    }

    /* Lock-free snapshot; all zero before the first StartRender. */
    Stats GetStats() const {
        mt_stats s = GetWorkStats();
        return {s.progress, s.time_elapsed, s.estimated_time_remaining,
                static_cast<std::size_t>(s.rays), static_cast<std::size_t>(s.samples)};
    }

    /* The full mt_stats (traversal work counters, kernel times); no reference counterpart. */
    mt_stats GetWorkStats() const {
        mt_stats s{};
        if (job_) Check(mt_render_stats(job_, &s), "mt_render_stats");
        return s;
    }

    /* A render was started and has ended (Tracer.cpp:823-828). */
    bool IsDone() const noexcept { return job_ && mt_render_is_done(job_); }

  private:
    std::shared_ptr<RenderBuffer> imageBuffer_;
    const Scene& scene_;
    mt_camera_desc cam_;
    mt_settings settings_;
    mt_job* job_ = nullptr;
};

}  // namespace mitrace_gpu

#endif
