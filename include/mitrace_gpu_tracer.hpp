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
 *                                     SaveToFilePNG / SaveToFileJPG / SaveToFileHDR
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
     * holds them.  JPG: baseline sequential JPEG (ITU-T T.81) of GetPixelsRGB8 in JFIF YCbCr,
     * see SaveToFileJPG.  A failure throws std::runtime_error("Failed to save image to file"),
     * the reference's message. */
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

    /* RenderBuffer.cpp:69-75 passes width * 3 (the row stride) where stb_image_write expects
     * the quality, which it clamps to 1..100: the file's quality is min(100, 3 * width), i.e.
     * 100 from 34 pixels wide on.  This encoder: JFIF YCbCr (T.871 conversion), one 8x8 block
     * per component and MCU (no chroma subsampling), the example quantisation tables of T.81
     * Annex K.1 scaled by the IJG quality rule (quality 100: every step 1), the example Huffman
     * tables of Annex K.3, float DCT-II with rounding to the nearest step, edge blocks padded
     * by repeating the last column and row. */
    void SaveToFileJPG(const std::string& filename) const {
        if (width_ == 0 || height_ == 0 || width_ > 65535 || height_ > 65535)
            throw std::runtime_error("Failed to save image to file");
        WriteFile(filename, EncodeJpeg(GetPixelsRGB8().get(), width_, height_,
                                       static_cast<int>(std::min<unsigned>(100u, width_ * 3u))));
    }

    float* Data() noexcept { return pixels_.data(); }
    const float* Data() const noexcept { return pixels_.data(); }

  private:
    // T.81 Annex K.3 Huffman tables: code counts per length 1..16, then the symbols in order
    struct HuffSpec {
        std::uint8_t counts[16];
        std::vector<std::uint8_t> symbols;
    };
    struct HuffCode {
        std::uint16_t code[256] = {};
        std::uint8_t len[256] = {};  // 0: symbol absent
    };
    static const HuffSpec& JpegTable(int which) {  // 0 DC luma, 1 AC luma, 2 DC chroma, 3 AC chroma
        static const HuffSpec specs[4] = {
            {{0, 1, 5, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0, 0, 0}, {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11}},
            {{0, 2, 1, 3, 3, 2, 4, 3, 5, 5, 4, 4, 0, 0, 1, 0x7d},
             {0x01, 0x02, 0x03, 0x00, 0x04, 0x11, 0x05, 0x12, 0x21, 0x31, 0x41, 0x06, 0x13, 0x51,
              0x61, 0x07, 0x22, 0x71, 0x14, 0x32, 0x81, 0x91, 0xa1, 0x08, 0x23, 0x42, 0xb1, 0xc1,
              0x15, 0x52, 0xd1, 0xf0, 0x24, 0x33, 0x62, 0x72, 0x82, 0x09, 0x0a, 0x16, 0x17, 0x18,
              0x19, 0x1a, 0x25, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x34, 0x35, 0x36, 0x37, 0x38, 0x39,
              0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56, 0x57,
              0x58, 0x59, 0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74, 0x75,
              0x76, 0x77, 0x78, 0x79, 0x7a, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89, 0x8a, 0x92,
              0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3, 0xa4, 0xa5, 0xa6, 0xa7,
              0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba, 0xc2, 0xc3,
              0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6, 0xd7, 0xd8,
              0xd9, 0xda, 0xe1, 0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf1, 0xf2,
              0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa}},
            {{0, 3, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 0, 0, 0, 0}, {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11}},
            {{0, 2, 1, 2, 4, 4, 3, 4, 7, 5, 4, 4, 0, 1, 2, 0x77},
             {0x00, 0x01, 0x02, 0x03, 0x11, 0x04, 0x05, 0x21, 0x31, 0x06, 0x12, 0x41, 0x51, 0x07,
              0x61, 0x71, 0x13, 0x22, 0x32, 0x81, 0x08, 0x14, 0x42, 0x91, 0xa1, 0xb1, 0xc1, 0x09,
              0x23, 0x33, 0x52, 0xf0, 0x15, 0x62, 0x72, 0xd1, 0x0a, 0x16, 0x24, 0x34, 0xe1, 0x25,
              0xf1, 0x17, 0x18, 0x19, 0x1a, 0x26, 0x27, 0x28, 0x29, 0x2a, 0x35, 0x36, 0x37, 0x38,
              0x39, 0x3a, 0x43, 0x44, 0x45, 0x46, 0x47, 0x48, 0x49, 0x4a, 0x53, 0x54, 0x55, 0x56,
              0x57, 0x58, 0x59, 0x5a, 0x63, 0x64, 0x65, 0x66, 0x67, 0x68, 0x69, 0x6a, 0x73, 0x74,
              0x75, 0x76, 0x77, 0x78, 0x79, 0x7a, 0x82, 0x83, 0x84, 0x85, 0x86, 0x87, 0x88, 0x89,
              0x8a, 0x92, 0x93, 0x94, 0x95, 0x96, 0x97, 0x98, 0x99, 0x9a, 0xa2, 0xa3, 0xa4, 0xa5,
              0xa6, 0xa7, 0xa8, 0xa9, 0xaa, 0xb2, 0xb3, 0xb4, 0xb5, 0xb6, 0xb7, 0xb8, 0xb9, 0xba,
              0xc2, 0xc3, 0xc4, 0xc5, 0xc6, 0xc7, 0xc8, 0xc9, 0xca, 0xd2, 0xd3, 0xd4, 0xd5, 0xd6,
              0xd7, 0xd8, 0xd9, 0xda, 0xe2, 0xe3, 0xe4, 0xe5, 0xe6, 0xe7, 0xe8, 0xe9, 0xea, 0xf2,
              0xf3, 0xf4, 0xf5, 0xf6, 0xf7, 0xf8, 0xf9, 0xfa}}};
        return specs[which];
    }
    // canonical codes from the counts (T.81 Annex C): consecutive codes within a length,
    // shifted left by one at each new length
    static HuffCode JpegCodes(const HuffSpec& s) {
        HuffCode h;
        unsigned code = 0;
        std::size_t k = 0;
        for (int l = 1; l <= 16; ++l, code <<= 1)
            for (int i = 0; i < s.counts[l - 1]; ++i, ++code, ++k) {
                h.code[s.symbols[k]] = static_cast<std::uint16_t>(code);
                h.len[s.symbols[k]] = static_cast<std::uint8_t>(l);
            }
        return h;
    }
    struct BitWriter {
        std::vector<std::uint8_t>& out;
        std::uint32_t acc = 0;
        int n = 0;
        void Put(std::uint32_t bits, int len) {  // most significant bit first
            for (int i = len - 1; i >= 0; --i) {
                acc = (acc << 1) | ((bits >> i) & 1u);
                if (++n == 8) {
                    out.push_back(static_cast<std::uint8_t>(acc));
                    if (acc == 0xff) out.push_back(0);  // byte stuffing (T.81 F.1.2.3)
                    acc = 0, n = 0;
                }
            }
        }
        void Flush() {  // pad the last byte with 1-bits
            while (n) Put(1, 1);
        }
    };
    static void PutSegment(std::vector<std::uint8_t>& f, std::uint8_t marker,
                           const std::vector<std::uint8_t>& body) {
        const std::size_t len = body.size() + 2;
        f.insert(f.end(), {0xff, marker, static_cast<std::uint8_t>(len >> 8),
                           static_cast<std::uint8_t>(len & 0xff)});
        f.insert(f.end(), body.begin(), body.end());
    }
    static std::vector<std::uint8_t> EncodeJpeg(const std::uint8_t* rgb, unsigned w, unsigned h,
                                                int quality) {
        // zig-zag scan: position k -> row-major index, walking the anti-diagonals alternately
        int zz[64];
        for (int k = 0, s = 0; s < 15; ++s)
            for (int i = 0; i <= s; ++i) {
                const int r = (s & 1) ? i : s - i, c = s - r;
                if (r < 8 && c < 8) zz[k++] = r * 8 + c;
            }
        static const std::uint8_t kLuma[64] = {
            16, 11, 10, 16, 24,  40,  51,  61,  12, 12, 14, 19, 26,  58,  60,  55,
            14, 13, 16, 24, 40,  57,  69,  56,  14, 17, 22, 29, 51,  87,  80,  62,
            18, 22, 37, 56, 68,  109, 103, 77,  24, 35, 55, 64, 81,  104, 113, 92,
            49, 64, 78, 87, 103, 121, 120, 101, 72, 92, 95, 98, 112, 100, 103, 99};
        static const std::uint8_t kChroma[64] = {
            17, 18, 24, 47, 99, 99, 99, 99, 18, 21, 26, 66, 99, 99, 99, 99,
            24, 26, 56, 99, 99, 99, 99, 99, 47, 66, 99, 99, 99, 99, 99, 99,
            99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99,
            99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99, 99};
        const int scale = quality < 50 ? 5000 / quality : 200 - 2 * quality;  // IJG rule
        std::uint8_t q[2][64];  // natural (row-major) order
        for (int i = 0; i < 64; ++i) {
            q[0][i] = static_cast<std::uint8_t>(std::clamp((kLuma[i] * scale + 50) / 100, 1, 255));
            q[1][i] = static_cast<std::uint8_t>(std::clamp((kChroma[i] * scale + 50) / 100, 1, 255));
        }
        std::vector<std::uint8_t> f = {0xff, 0xd8};  // SOI
        PutSegment(f, 0xe0, {'J', 'F', 'I', 'F', 0, 1, 1, 0, 0, 1, 0, 1, 0, 0});
        for (int t = 0; t < 2; ++t) {
            std::vector<std::uint8_t> dqt = {static_cast<std::uint8_t>(t)};
            for (int k = 0; k < 64; ++k) dqt.push_back(q[t][zz[k]]);
            PutSegment(f, 0xdb, dqt);
        }
        PutSegment(f, 0xc0, {8, static_cast<std::uint8_t>(h >> 8), static_cast<std::uint8_t>(h & 0xff),
                             static_cast<std::uint8_t>(w >> 8), static_cast<std::uint8_t>(w & 0xff), 3,
                             1, 0x11, 0, 2, 0x11, 1, 3, 0x11, 1});
        const std::uint8_t tclass[4] = {0x00, 0x10, 0x01, 0x11};
        for (int t = 0; t < 4; ++t) {
            const HuffSpec& s = JpegTable(t);
            std::vector<std::uint8_t> dht = {tclass[t]};
            dht.insert(dht.end(), s.counts, s.counts + 16);
            dht.insert(dht.end(), s.symbols.begin(), s.symbols.end());
            PutSegment(f, 0xc4, dht);
        }
        PutSegment(f, 0xda, {3, 1, 0x00, 2, 0x11, 3, 0x11, 0, 63, 0});
        const HuffCode codes[4] = {JpegCodes(JpegTable(0)), JpegCodes(JpegTable(1)),
                                   JpegCodes(JpegTable(2)), JpegCodes(JpegTable(3))};
        double basis[8][8];  // C(u) cos((2x + 1) u pi / 16) / 2
        for (int u = 0; u < 8; ++u)
            for (int x = 0; x < 8; ++x)
                basis[u][x] = (u ? 0.5 : 0.5 / std::sqrt(2.0)) *
                              std::cos((2 * x + 1) * u * 3.14159265358979323846 / 16.0);
        BitWriter bw{f};
        auto put_sym = [&](const HuffCode& hc, int sym) {
            if (!hc.len[sym]) throw std::runtime_error("Failed to save image to file");
            bw.Put(hc.code[sym], hc.len[sym]);
        };
        auto magnitude = [](int v, int& cat) {  // category and its appended bits (T.81 F.1.2.1)
            const int a = v < 0 ? -v : v;
            cat = 0;
            while ((a >> cat) > 0) ++cat;
            return static_cast<std::uint32_t>(v < 0 ? v + (1 << cat) - 1 : v);
        };
        int pred[3] = {0, 0, 0};
        for (unsigned by = 0; by < h; by += 8)
            for (unsigned bx = 0; bx < w; bx += 8) {
                double blk[3][64];
                for (int y = 0; y < 8; ++y)
                    for (int x = 0; x < 8; ++x) {
                        const unsigned sx = std::min(bx + x, w - 1), sy = std::min(by + y, h - 1);
                        const std::uint8_t* p = rgb + (std::size_t(sy) * w + sx) * 3;
                        const double r = p[0], g = p[1], b = p[2];
                        blk[0][y * 8 + x] = 0.299 * r + 0.587 * g + 0.114 * b - 128.0;
                        blk[1][y * 8 + x] = -0.168735892 * r - 0.331264108 * g + 0.5 * b;
                        blk[2][y * 8 + x] = 0.5 * r - 0.418687589 * g - 0.081312411 * b;
                    }
                for (int c = 0; c < 3; ++c) {
                    double tmp[64];
                    int coef[64];
                    for (int y = 0; y < 8; ++y)  // rows, then columns
                        for (int u = 0; u < 8; ++u) {
                            double s = 0;
                            for (int x = 0; x < 8; ++x) s += basis[u][x] * blk[c][y * 8 + x];
                            tmp[y * 8 + u] = s;
                        }
                    for (int v = 0; v < 8; ++v)
                        for (int u = 0; u < 8; ++u) {
                            double s = 0;
                            for (int y = 0; y < 8; ++y) s += basis[v][y] * tmp[y * 8 + u];
                            coef[v * 8 + u] = static_cast<int>(std::lround(s / q[c ? 1 : 0][v * 8 + u]));
                        }
                    const HuffCode& dc = codes[c ? 2 : 0];
                    const HuffCode& ac = codes[c ? 3 : 1];
                    int cat = 0;
                    std::uint32_t bits = magnitude(coef[0] - pred[c], cat);
                    pred[c] = coef[0];
                    put_sym(dc, cat);
                    bw.Put(bits, cat);
                    int run = 0;
                    for (int k = 1; k < 64; ++k) {
                        const int v = coef[zz[k]];
                        if (v == 0) {
                            ++run;
                            continue;
                        }
                        for (; run >= 16; run -= 16) put_sym(ac, 0xf0);  // ZRL
                        bits = magnitude(v, cat);
                        put_sym(ac, (run << 4) | cat);
                        bw.Put(bits, cat);
                        run = 0;
                    }
                    if (run) put_sym(ac, 0x00);  // EOB
                }
            }
        bw.Flush();
        f.insert(f.end(), {0xff, 0xd9});  // EOI
        return f;
    }

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
