/*
 * mitrace_gpu_loader.hpp -- C++20 scene loading for the GPU Tracer: the reference loader's
 * glTF subset and semantics (src/Loader/GLTF.cpp, src/Loader/Types.hpp) plus Scene::Scene's
 * assembly (src/Scene/Scene.cpp:9-59), producing the C-ABI's mt_scene_desc.  Header-only,
 * no dependencies beyond the C++ standard library: it carries its own JSON reader, zlib
 * inflate and PNG decoder (the reference reads images through stb_image) and a Radiance .hdr
 * reader for float environments.
 *
 *   this header                          reference
 *   -----------------------------------  ------------------------------------------------------
 *   LoadGltf(path, scene, camera)        Loader::GLTF(path).LoadScene(scene) + Scene::Scene +
 *                                          SetCamera (main.cpp:100-107, GLTF.cpp:681-939)
 *   LoadedScene::SetEnvironment(path)    --hdri-override (main.cpp:108-112, Texture.cpp:35-50)
 *   LoadedScene::ToDesc()                the scene as mt_scene_create takes it
 *   LoadImageRgb8(path)                  Image::Image -> stbi_load(path, ..., 3) (Image.cpp:77-104):
 *                                          8-bit RGB, alpha dropped, 16-bit samples -> high byte
 *   CliOptions / ToSettings(options)     Config::ConfigOptions (src/CLI/Config.hpp:43-80) -> mt_settings
 *
 * Accepted glTF subset, as the reference's: buffers in external files (data URIs rejected),
 * triangle primitives (mode 4) with indices (UINT16/UINT32), tightly packed FLOAT accessors
 * (byteStride ignored), POSITION and NORMAL required, TEXCOORD_0 optional; pbrMetallicRoughness
 * materials (a primitive without one gets Loader::Material{}: metallic and roughness 0.5,
 * occlusion 0), KHR_materials_emissive_strength, KHR_texture_transform offset/scale;
 * KHR_lights_punctual point/spot/directional lights and EXT_light_attributes area lights;
 * perspective and orthographic cameras; node trees with matrix or T*R*S, in depth-first
 * order.  Images: PNG (8- and 16-bit, all colour types, Adam7 interlace); other image formats
 * are rejected with an error.  Errors throw std::runtime_error with the reference's message
 * where it has one.
 */
#ifndef MITRACE_GPU_LOADER_HPP
#define MITRACE_GPU_LOADER_HPP

#include <algorithm>
#include <array>
#include <charconv>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "mitrace_gpu.h"

namespace mitrace_gpu {

// ------------------------------------------------------------------------------------ JSON
namespace json {

class Value {
  public:
    enum class Kind { Null, Bool, Number, String, Array, Object };

    Kind kind() const noexcept { return kind_; }
    bool Has(const std::string& key) const noexcept { return Find(key) != nullptr; }
    const Value& operator[](const std::string& key) const {
        const Value* v = Find(key);
        if (!v) throw std::runtime_error("glTF: missing key '" + key + "'");
        return *v;
    }
    const Value& operator[](std::size_t i) const {
        if (kind_ != Kind::Array || i >= arr_.size()) throw std::runtime_error("glTF: index out of range");
        return arr_[i];
    }
    std::size_t Size() const noexcept { return kind_ == Kind::Array ? arr_.size() : obj_.size(); }
    double Num() const {
        if (kind_ != Kind::Number) throw std::runtime_error("glTF: number expected");
        return num_;
    }
    std::size_t Index() const {
        const double d = Num();
        if (!(d >= 0.0) || d != std::floor(d) || d > 4294967295.0) throw std::runtime_error("glTF: index expected");
        return static_cast<std::size_t>(d);
    }
    const std::string& Str() const {
        if (kind_ != Kind::String) throw std::runtime_error("glTF: string expected");
        return str_;
    }
    // the value of `key`, or `def` when it is absent
    double NumOr(const std::string& key, double def) const {
        const Value* v = Find(key);
        return v ? v->Num() : def;
    }
    const Value* Find(const std::string& key) const noexcept {
        if (kind_ != Kind::Object) return nullptr;
        for (const auto& [k, v] : obj_)
            if (k == key) return &v;
        return nullptr;
    }

    static Value Parse(const std::string& text) {
        std::size_t pos = 0;
        Value v = ParseValue(text, pos, 0);
        SkipWs(text, pos);
        if (pos != text.size()) throw std::runtime_error("glTF: trailing characters after JSON");
        return v;
    }

  private:
    Kind kind_ = Kind::Null;
    bool b_ = false;
    double num_ = 0.0;
    std::string str_;
    std::vector<Value> arr_;
    std::vector<std::pair<std::string, Value>> obj_;

    static void SkipWs(const std::string& s, std::size_t& p) {
        while (p < s.size() && (s[p] == ' ' || s[p] == '\t' || s[p] == '\n' || s[p] == '\r')) ++p;
    }
    [[noreturn]] static void Fail(const char* what) {
        throw std::runtime_error(std::string("glTF: invalid JSON (") + what + ")");
    }
    static void PutUtf8(std::string& out, uint32_t cp) {
        if (cp < 0x80) {
            out += static_cast<char>(cp);
        } else if (cp < 0x800) {
            out += static_cast<char>(0xC0 | (cp >> 6));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        } else if (cp < 0x10000) {
            out += static_cast<char>(0xE0 | (cp >> 12));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        } else {
            out += static_cast<char>(0xF0 | (cp >> 18));
            out += static_cast<char>(0x80 | ((cp >> 12) & 0x3F));
            out += static_cast<char>(0x80 | ((cp >> 6) & 0x3F));
            out += static_cast<char>(0x80 | (cp & 0x3F));
        }
    }
    static uint32_t Hex4(const std::string& s, std::size_t& p) {
        if (p + 4 > s.size()) Fail("short \\u escape");
        uint32_t v = 0;
        for (int i = 0; i < 4; ++i) {
            const char c = s[p++];
            v <<= 4;
            if (c >= '0' && c <= '9') v |= static_cast<uint32_t>(c - '0');
            else if (c >= 'a' && c <= 'f') v |= static_cast<uint32_t>(c - 'a' + 10);
            else if (c >= 'A' && c <= 'F') v |= static_cast<uint32_t>(c - 'A' + 10);
            else Fail("bad \\u escape");
        }
        return v;
    }
    static std::string ParseString(const std::string& s, std::size_t& p) {
        if (s[p] != '"') Fail("string expected");
        ++p;
        std::string out;
        while (true) {
            if (p >= s.size()) Fail("unterminated string");
            const char c = s[p++];
            if (c == '"') break;
            if (c != '\\') {
                out += c;
                continue;
            }
            if (p >= s.size()) Fail("unterminated escape");
            const char e = s[p++];
            switch (e) {
                case '"': out += '"'; break;
                case '\\': out += '\\'; break;
                case '/': out += '/'; break;
                case 'b': out += '\b'; break;
                case 'f': out += '\f'; break;
                case 'n': out += '\n'; break;
                case 'r': out += '\r'; break;
                case 't': out += '\t'; break;
                case 'u': {
                    uint32_t cp = Hex4(s, p);
                    if (cp >= 0xD800 && cp < 0xDC00 && p + 6 <= s.size() && s[p] == '\\' && s[p + 1] == 'u') {
                        p += 2;
                        const uint32_t lo = Hex4(s, p);
                        cp = 0x10000 + ((cp - 0xD800) << 10) + (lo - 0xDC00);
                    }
                    PutUtf8(out, cp);
                    break;
                }
                default: Fail("bad escape");
            }
        }
        return out;
    }
    static Value ParseValue(const std::string& s, std::size_t& p, int depth) {
        if (depth > 512) Fail("nesting too deep");
        SkipWs(s, p);
        if (p >= s.size()) Fail("unexpected end");
        Value v;
        const char c = s[p];
        if (c == '{') {
            v.kind_ = Kind::Object;
            ++p;
            SkipWs(s, p);
            if (p < s.size() && s[p] == '}') {
                ++p;
                return v;
            }
            while (true) {
                SkipWs(s, p);
                if (p >= s.size()) Fail("unterminated object");
                std::string key = ParseString(s, p);
                SkipWs(s, p);
                if (p >= s.size() || s[p] != ':') Fail("':' expected");
                ++p;
                Value item = ParseValue(s, p, depth + 1);
                // a repeated key keeps its last value, as Python's json does
                bool replaced = false;
                for (auto& kv : v.obj_)
                    if (kv.first == key) {
                        kv.second = std::move(item);
                        replaced = true;
                        break;
                    }
                if (!replaced) v.obj_.emplace_back(std::move(key), std::move(item));
                SkipWs(s, p);
                if (p < s.size() && s[p] == ',') {
                    ++p;
                    continue;
                }
                if (p < s.size() && s[p] == '}') {
                    ++p;
                    return v;
                }
                Fail("',' or '}' expected");
            }
        }
        if (c == '[') {
            v.kind_ = Kind::Array;
            ++p;
            SkipWs(s, p);
            if (p < s.size() && s[p] == ']') {
                ++p;
                return v;
            }
            while (true) {
                v.arr_.push_back(ParseValue(s, p, depth + 1));
                SkipWs(s, p);
                if (p < s.size() && s[p] == ',') {
                    ++p;
                    continue;
                }
                if (p < s.size() && s[p] == ']') {
                    ++p;
                    return v;
                }
                Fail("',' or ']' expected");
            }
        }
        if (c == '"') {
            v.kind_ = Kind::String;
            v.str_ = ParseString(s, p);
            return v;
        }
        if (s.compare(p, 4, "true") == 0) {
            v.kind_ = Kind::Bool;
            v.b_ = true;
            p += 4;
            return v;
        }
        if (s.compare(p, 5, "false") == 0) {
            v.kind_ = Kind::Bool;
            p += 5;
            return v;
        }
        if (s.compare(p, 4, "null") == 0) {
            p += 4;
            return v;
        }
        // number: from_chars rounds correctly and does not depend on the locale
        const char* b = s.data() + p;
        const char* e = s.data() + s.size();
        const char* q = b;
        if (q < e && *q == '-') ++q;
        double d = 0.0;
        const auto r = std::from_chars(q, e, d);
        if (r.ec != std::errc() || r.ptr == q) Fail("number expected");
        v.kind_ = Kind::Number;
        v.num_ = (*b == '-') ? -d : d;
        p = static_cast<std::size_t>(r.ptr - s.data());
        return v;
    }
};

}  // namespace json

// ------------------------------------------------------------------------------------ images
namespace image {

inline std::vector<uint8_t> ReadFile(const std::filesystem::path& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("Failed to open file: " + path.string());
    return std::vector<uint8_t>(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
}

// DEFLATE (RFC 1951) decoder: stored, fixed-Huffman and dynamic-Huffman blocks
class Inflater {
  public:
    Inflater(const uint8_t* data, std::size_t n) : d_(data), n_(n) {}

    std::vector<uint8_t> Run() {
        std::vector<uint8_t> out;
        bool last = false;
        while (!last) {
            last = Bits(1) != 0;
            const uint32_t type = Bits(2);
            if (type == 0) {
                bitbuf_ = 0;  // to the byte boundary
                bitcnt_ = 0;
                if (pos_ + 4 > n_) Fail();
                const uint32_t len = d_[pos_] | (d_[pos_ + 1] << 8);
                const uint32_t nlen = d_[pos_ + 2] | (d_[pos_ + 3] << 8);
                pos_ += 4;
                if ((len ^ 0xFFFFu) != nlen || pos_ + len > n_) Fail();
                out.insert(out.end(), d_ + pos_, d_ + pos_ + len);
                pos_ += len;
            } else if (type == 1 || type == 2) {
                Huffman lit, dist;
                if (type == 1) {
                    uint8_t lens[288 + 32];
                    for (int i = 0; i < 144; ++i) lens[i] = 8;
                    for (int i = 144; i < 256; ++i) lens[i] = 9;
                    for (int i = 256; i < 280; ++i) lens[i] = 7;
                    for (int i = 280; i < 288; ++i) lens[i] = 8;
                    for (int i = 0; i < 32; ++i) lens[288 + i] = 5;
                    lit.Build(lens, 288);
                    dist.Build(lens + 288, 32);
                } else {
                    ReadDynamic(lit, dist);
                }
                Codes(lit, dist, out);
            } else {
                Fail();
            }
        }
        return out;
    }

  private:
    struct Huffman {
        uint16_t count[16] = {};
        std::vector<uint16_t> symbol;
        void Build(const uint8_t* lens, int n) {
            std::fill(std::begin(count), std::end(count), uint16_t(0));
            for (int i = 0; i < n; ++i) count[lens[i]]++;
            count[0] = 0;
            uint16_t offs[16] = {};
            for (int l = 1; l < 16; ++l) offs[l] = static_cast<uint16_t>(offs[l - 1] + count[l - 1]);
            symbol.assign(static_cast<std::size_t>(n), 0);
            for (int i = 0; i < n; ++i)
                if (lens[i]) symbol[offs[lens[i]]++] = static_cast<uint16_t>(i);
        }
    };

    const uint8_t* d_;
    std::size_t n_, pos_ = 0;
    uint32_t bitbuf_ = 0;
    int bitcnt_ = 0;

    [[noreturn]] static void Fail() { throw std::runtime_error("PNG: corrupt deflate stream"); }
    uint32_t Bits(int need) {
        while (bitcnt_ < need) {
            if (pos_ >= n_) Fail();
            bitbuf_ |= static_cast<uint32_t>(d_[pos_++]) << bitcnt_;
            bitcnt_ += 8;
        }
        const uint32_t v = bitbuf_ & ((1u << need) - 1u);
        bitbuf_ >>= need;
        bitcnt_ -= need;
        return v;
    }
    // canonical code, read one bit at a time, most significant code bit first
    int Decode(const Huffman& h) {
        int code = 0, first = 0, index = 0;
        for (int len = 1; len < 16; ++len) {
            code |= static_cast<int>(Bits(1));
            const int cnt = h.count[len];
            if (code - cnt < first) return h.symbol[static_cast<std::size_t>(index + (code - first))];
            index += cnt;
            first = (first + cnt) << 1;
            code <<= 1;
        }
        Fail();
    }
    void ReadDynamic(Huffman& lit, Huffman& dist) {
        static const uint8_t kOrder[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};
        const int nlen = static_cast<int>(Bits(5)) + 257, ndist = static_cast<int>(Bits(5)) + 1;
        const int ncode = static_cast<int>(Bits(4)) + 4;
        if (nlen > 286 || ndist > 30) Fail();
        uint8_t lens[320] = {};
        for (int i = 0; i < ncode; ++i) lens[kOrder[i]] = static_cast<uint8_t>(Bits(3));
        Huffman code;
        code.Build(lens, 19);
        int i = 0;
        uint8_t all[320] = {};
        while (i < nlen + ndist) {
            const int sym = Decode(code);
            if (sym < 16) {
                all[i++] = static_cast<uint8_t>(sym);
                continue;
            }
            int rep = 0;
            uint8_t val = 0;
            if (sym == 16) {
                if (i == 0) Fail();
                val = all[i - 1];
                rep = 3 + static_cast<int>(Bits(2));
            } else if (sym == 17) {
                rep = 3 + static_cast<int>(Bits(3));
            } else {
                rep = 11 + static_cast<int>(Bits(7));
            }
            if (i + rep > nlen + ndist) Fail();
            while (rep--) all[i++] = val;
        }
        lit.Build(all, nlen);
        dist.Build(all + nlen, ndist);
    }
    void Codes(const Huffman& lit, const Huffman& dist, std::vector<uint8_t>& out) {
        static const uint16_t kLenBase[29] = {3, 4, 5, 6, 7, 8, 9, 10, 11, 13, 15, 17, 19, 23, 27, 31,
                                              35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
        static const uint8_t kLenExtra[29] = {0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 2, 2, 2, 2,
                                              3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5, 0};
        static const uint16_t kDistBase[30] = {1, 2, 3, 4, 5, 7, 9, 13, 17, 25, 33, 49, 65, 97, 129,
                                               193, 257, 385, 513, 769, 1025, 1537, 2049, 3073, 4097,
                                               6145, 8193, 12289, 16385, 24577};
        static const uint8_t kDistExtra[30] = {0, 0, 0, 0, 1, 1, 2, 2, 3, 3, 4, 4, 5, 5, 6,
                                               6, 7, 7, 8, 8, 9, 9, 10, 10, 11, 11, 12, 12, 13, 13};
        while (true) {
            int sym = Decode(lit);
            if (sym < 256) {
                out.push_back(static_cast<uint8_t>(sym));
                continue;
            }
            if (sym == 256) return;
            sym -= 257;
            if (sym >= 29) Fail();
            const std::size_t len = kLenBase[sym] + Bits(kLenExtra[sym]);
            const int ds = Decode(dist);
            if (ds >= 30) Fail();
            const std::size_t back = kDistBase[ds] + Bits(kDistExtra[ds]);
            if (back > out.size()) Fail();
            const std::size_t from = out.size() - back;
            for (std::size_t k = 0; k < len; ++k) out.push_back(out[from + k]);
        }
    }
};

// zlib stream (RFC 1950): header, DEFLATE data, Adler-32 of the output
inline std::vector<uint8_t> Inflate(const uint8_t* data, std::size_t n) {
    if (n < 6 || (data[0] & 0x0F) != 8 || ((data[0] << 8) | data[1]) % 31 != 0 || (data[1] & 0x20))
        throw std::runtime_error("PNG: bad zlib header");
    std::vector<uint8_t> out = Inflater(data + 2, n - 6).Run();
    uint32_t a = 1, b = 0;
    for (uint8_t c : out) {
        a = (a + c) % 65521u;
        b = (b + a) % 65521u;
    }
    const uint8_t* t = data + n - 4;
    const uint32_t want = (uint32_t(t[0]) << 24) | (uint32_t(t[1]) << 16) | (uint32_t(t[2]) << 8) | t[3];
    if (((b << 16) | a) != want) throw std::runtime_error("PNG: zlib checksum mismatch");
    return out;
}

struct Rgb8 {
    uint32_t width = 0, height = 0;
    std::vector<uint8_t> rgb;  // row-major, 3 bytes per texel
};

// PNG (ISO/IEC 15948) to 8-bit RGB, as stbi_load(..., 3) returns it: grey replicated, alpha
// dropped, palette looked up, 16-bit samples reduced to their high byte, grey of 1/2/4 bits
// scaled to 0..255
inline Rgb8 DecodePng(const std::vector<uint8_t>& f) {
    static const uint8_t kSig[8] = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1a, '\n'};
    if (f.size() < 8 || std::memcmp(f.data(), kSig, 8) != 0) throw std::runtime_error("not a PNG file");
    auto be32 = [&](std::size_t p) {
        return (uint32_t(f[p]) << 24) | (uint32_t(f[p + 1]) << 16) | (uint32_t(f[p + 2]) << 8) | f[p + 3];
    };
    uint32_t w = 0, h = 0, depth = 0, ctype = 0, interlace = 0;
    std::vector<uint8_t> idat, plte;
    std::size_t p = 8;
    bool seen_ihdr = false, seen_iend = false;
    while (p + 12 <= f.size() && !seen_iend) {
        const uint32_t len = be32(p);
        if (p + 12 + std::size_t(len) > f.size()) throw std::runtime_error("PNG: truncated chunk");
        const char* type = reinterpret_cast<const char*>(&f[p + 4]);
        uint32_t crc = 0xffffffffu;  // CRC-32 (ISO 3309) over type and data
        for (std::size_t i = p + 4; i < p + 8 + len; ++i) {
            crc ^= f[i];
            for (int k = 0; k < 8; ++k) crc = (crc >> 1) ^ (0xedb88320u & (0u - (crc & 1u)));
        }
        if (~crc != be32(p + 8 + len)) throw std::runtime_error("PNG: chunk CRC mismatch");
        const uint8_t* body = &f[p + 8];
        if (std::memcmp(type, "IHDR", 4) == 0) {
            if (len != 13) throw std::runtime_error("PNG: bad IHDR");
            w = be32(p + 8);
            h = be32(p + 12);
            depth = body[8];
            ctype = body[9];
            if (body[10] != 0 || body[11] != 0) throw std::runtime_error("PNG: unknown compression or filter");
            interlace = body[12];
            seen_ihdr = true;
        } else if (std::memcmp(type, "PLTE", 4) == 0) {
            plte.assign(body, body + len);
        } else if (std::memcmp(type, "IDAT", 4) == 0) {
            idat.insert(idat.end(), body, body + len);
        } else if (std::memcmp(type, "IEND", 4) == 0) {
            seen_iend = true;
        } else if (!(type[0] & 0x20)) {
            throw std::runtime_error("PNG: unknown critical chunk");
        }
        p += 12 + std::size_t(len);
    }
    if (!seen_ihdr || idat.empty() || w == 0 || h == 0 || interlace > 1)
        throw std::runtime_error("PNG: missing or bad image data");
    int channels = 0;
    switch (ctype) {
        case 0: channels = 1; break;
        case 2: channels = 3; break;
        case 3: channels = 1; break;
        case 4: channels = 2; break;
        case 6: channels = 4; break;
        default: throw std::runtime_error("PNG: bad colour type");
    }
    const bool depth_ok = depth == 8 || (depth == 16 && ctype != 3) ||
                          ((depth == 1 || depth == 2 || depth == 4) && (ctype == 0 || ctype == 3));
    if (!depth_ok) throw std::runtime_error("PNG: bad bit depth");
    if (ctype == 3 && (plte.empty() || plte.size() % 3)) throw std::runtime_error("PNG: missing palette");
    if (uint64_t(w) * h > (uint64_t(1) << 30)) throw std::runtime_error("PNG: image too large");
    const std::vector<uint8_t> raw = Inflate(idat.data(), idat.size());
    const uint32_t bits = depth * static_cast<uint32_t>(channels);
    const uint32_t bpp = std::max<uint32_t>(1u, bits / 8u);  // filter unit in bytes
    Rgb8 img;
    img.width = w;
    img.height = h;
    img.rgb.assign(std::size_t(w) * h * 3, 0);
    // one sample (channel c of pixel x) of an unfiltered row
    auto sample = [&](const uint8_t* row, uint32_t x, int c) -> uint32_t {
        if (depth == 8) return row[std::size_t(x) * channels + c];
        if (depth == 16) return row[(std::size_t(x) * channels + c) * 2];  // high byte
        const uint32_t bit = x * depth;
        return (row[bit >> 3] >> (8 - depth - (bit & 7))) & ((1u << depth) - 1u);
    };
    auto put = [&](const uint8_t* row, uint32_t x, uint32_t ox, uint32_t oy) {
        uint8_t* o = &img.rgb[(std::size_t(oy) * w + ox) * 3];
        if (ctype == 3) {
            const uint32_t idx = sample(row, x, 0);
            if (std::size_t(idx) * 3 + 2 >= plte.size()) throw std::runtime_error("PNG: palette index out of range");
            o[0] = plte[idx * 3];
            o[1] = plte[idx * 3 + 1];
            o[2] = plte[idx * 3 + 2];
        } else if (ctype == 0 || ctype == 4) {
            uint32_t g = sample(row, x, 0);
            if (depth < 8) g = g * (255u / ((1u << depth) - 1u));
            o[0] = o[1] = o[2] = static_cast<uint8_t>(g);
        } else {
            for (int c = 0; c < 3; ++c) o[c] = static_cast<uint8_t>(sample(row, x, c));
        }
    };
    // undoes the filters of a (sub)image of pw x ph pixels starting at raw[pos]; calls put
    // for each pixel with its position in the pass
    std::size_t pos = 0;
    auto pass = [&](uint32_t pw, uint32_t ph, uint32_t x0, uint32_t y0, uint32_t dx, uint32_t dy) {
        if (pw == 0 || ph == 0) return;
        const std::size_t stride = (std::size_t(pw) * bits + 7) / 8;
        std::vector<uint8_t> prev(stride, 0), cur(stride);
        for (uint32_t y = 0; y < ph; ++y) {
            if (pos + 1 + stride > raw.size()) throw std::runtime_error("PNG: image data too short");
            const uint8_t ft = raw[pos++];
            for (std::size_t i = 0; i < stride; ++i) {
                const int a = i >= bpp ? cur[i - bpp] : 0, b = prev[i], c = i >= bpp ? prev[i - bpp] : 0;
                int pred = 0;
                switch (ft) {
                    case 0: pred = 0; break;
                    case 1: pred = a; break;
                    case 2: pred = b; break;
                    case 3: pred = (a + b) >> 1; break;
                    case 4: {
                        const int pp = a + b - c, pa = std::abs(pp - a), pb = std::abs(pp - b), pc = std::abs(pp - c);
                        pred = (pa <= pb && pa <= pc) ? a : (pb <= pc ? b : c);
                        break;
                    }
                    default: throw std::runtime_error("PNG: bad filter type");
                }
                cur[i] = static_cast<uint8_t>(raw[pos + i] + pred);
            }
            pos += stride;
            for (uint32_t x = 0; x < pw; ++x) put(cur.data(), x, x0 + x * dx, y0 + y * dy);
            std::swap(prev, cur);
        }
    };
    if (interlace == 0) {
        pass(w, h, 0, 0, 1, 1);
    } else {  // Adam7: seven passes over 8x8 blocks
        static const uint32_t kX0[7] = {0, 4, 0, 2, 0, 1, 0}, kY0[7] = {0, 0, 4, 0, 2, 0, 1};
        static const uint32_t kDx[7] = {8, 8, 4, 4, 2, 2, 1}, kDy[7] = {8, 8, 8, 4, 4, 2, 2};
        for (int k = 0; k < 7; ++k) {
            const uint32_t pw = w > kX0[k] ? (w - kX0[k] + kDx[k] - 1) / kDx[k] : 0;
            const uint32_t ph = h > kY0[k] ? (h - kY0[k] + kDy[k] - 1) / kDy[k] : 0;
            pass(pw, ph, kX0[k], kY0[k], kDx[k], kDy[k]);
        }
    }
    return img;
}

// Image::Image (Image.cpp:77-104): float images are rejected with the reference's message
inline Rgb8 LoadImageRgb8(const std::filesystem::path& path) {
    std::string ext = path.extension().string();
    std::transform(ext.begin(), ext.end(), ext.begin(), [](unsigned char c) { return std::tolower(c); });
    if (ext == ".hdr" || ext == ".exr") throw std::runtime_error("Cannot load float image");
    const std::vector<uint8_t> f = ReadFile(path);
    if (f.size() >= 8 && f[0] == 0x89 && f[1] == 'P') return DecodePng(f);
    throw std::runtime_error("Unsupported image format (PNG only): " + path.string());
}

struct RgbF {
    uint32_t width = 0, height = 0;
    std::vector<float> rgb;
};

// Radiance RGBE (.hdr): '-Y H +X W' orientation, flat or new-style run-length scanlines; a
// texel (m, e) decodes to m * 2^(e - 136), 0 when e = 0
inline RgbF LoadHdr(const std::filesystem::path& path) {
    const std::vector<uint8_t> d = ReadFile(path);
    auto starts = [&](const char* s) { return d.size() >= std::strlen(s) && std::memcmp(d.data(), s, std::strlen(s)) == 0; };
    if (!starts("#?RADIANCE") && !starts("#?RGBE")) throw std::runtime_error("not a Radiance HDR file");
    std::size_t p = 0;
    bool fmt_ok = false;
    auto line = [&]() {
        std::string s;
        while (p < d.size() && d[p] != '\n') s += static_cast<char>(d[p++]);
        if (p >= d.size()) throw std::runtime_error("HDR: truncated header");
        ++p;
        return s;
    };
    while (true) {
        const std::string s = line();
        if (s.empty()) break;
        if (s == "FORMAT=32-bit_rle_rgbe") fmt_ok = true;
    }
    if (!fmt_ok) throw std::runtime_error("HDR: unsupported format");
    const std::string res = line();
    int h = 0, w = 0;
    if (std::sscanf(res.c_str(), "-Y %d +X %d", &h, &w) != 2 || h <= 0 || w <= 0)
        throw std::runtime_error("HDR: unsupported orientation");
    RgbF img;
    img.width = static_cast<uint32_t>(w);
    img.height = static_cast<uint32_t>(h);
    img.rgb.assign(std::size_t(w) * h * 3, 0.0f);
    std::vector<uint8_t> scan(std::size_t(w) * 4);
    for (int y = 0; y < h; ++y) {
        if (w >= 8 && w < 32768 && p + 4 <= d.size() && d[p] == 2 && d[p + 1] == 2 && !(d[p + 2] & 0x80)) {
            if (((d[p + 2] << 8) | d[p + 3]) != w) throw std::runtime_error("HDR: bad scanline width");
            p += 4;
            for (int c = 0; c < 4; ++c) {  // each component run-length coded in turn
                int x = 0;
                while (x < w) {
                    if (p >= d.size()) throw std::runtime_error("HDR: truncated data");
                    int n = d[p++];
                    if (n > 128) {
                        n -= 128;
                        if (p >= d.size() || x + n > w) throw std::runtime_error("HDR: bad run");
                        const uint8_t v = d[p++];
                        while (n--) scan[std::size_t(x++) * 4 + c] = v;
                    } else {
                        if (n == 0 || x + n > w || p + n > d.size()) throw std::runtime_error("HDR: bad run");
                        while (n--) scan[std::size_t(x++) * 4 + c] = d[p++];
                    }
                }
            }
        } else {
            if (p + scan.size() > d.size()) throw std::runtime_error("HDR: truncated data");
            std::memcpy(scan.data(), &d[p], scan.size());
            p += scan.size();
        }
        for (int x = 0; x < w; ++x) {
            const uint8_t* t = &scan[std::size_t(x) * 4];
            float* o = &img.rgb[(std::size_t(y) * w + x) * 3];
            if (t[3] == 0) {
                o[0] = o[1] = o[2] = 0.0f;
            } else {
                const float s = std::ldexp(1.0f, int(t[3]) - 136);
                o[0] = t[0] * s;
                o[1] = t[1] * s;
                o[2] = t[2] * s;
            }
        }
    }
    return img;
}

}  // namespace image

// ------------------------------------------------------------------------------------ scene
// A loaded scene: the storage behind an mt_scene_desc (ToDesc points into it, so it must
// outlive the scene creation).
struct LoadedScene {
    struct Mesh {
        std::string name;
        std::vector<float> positions, normals, texcoord0;  // per corner: 3, 3, 2 floats
    };
    std::vector<Mesh> meshes;
    std::vector<image::Rgb8> images;
    std::vector<mt_material_desc> materials;
    std::vector<mt_instance_desc> instances;
    std::vector<mt_light_desc> lights;
    mt_camera_desc camera{};
    // environment: RGB8 or float texels (empty = none)
    std::vector<uint8_t> env_rgb8;
    std::vector<float> env_rgbf;
    uint32_t env_width = 0, env_height = 0;
    mt_bvh_options bvh{0, 32, 32, 0, 65536, 0};

    /* --hdri-override (main.cpp:108-112): an equirect from an 8-bit image (as the reference),
     * or float RGB from a Radiance .hdr; always sampled bilinearly (Texture.cpp:35-50) */
    void SetEnvironment(const std::filesystem::path& path) {
        std::string ext = path.extension().string();
        std::transform(ext.begin(), ext.end(), ext.begin(), [](unsigned char c) { return std::tolower(c); });
        env_rgb8.clear();
        env_rgbf.clear();
        if (ext == ".hdr") {
            image::RgbF f = image::LoadHdr(path);
            env_width = f.width;
            env_height = f.height;
            env_rgbf = std::move(f.rgb);
        } else {
            image::Rgb8 b = image::LoadImageRgb8(path);
            env_width = b.width;
            env_height = b.height;
            env_rgb8 = std::move(b.rgb);
        }
    }

    /* the C-ABI description; valid while this object lives and is not modified */
    mt_scene_desc ToDesc() const {
        mesh_descs_.clear();
        for (const Mesh& m : meshes)
            mesh_descs_.push_back({static_cast<uint32_t>(m.positions.size() / 9), m.positions.data(),
                                   m.normals.data(), m.texcoord0.empty() ? nullptr : m.texcoord0.data()});
        image_descs_.clear();
        for (const image::Rgb8& im : images) image_descs_.push_back({im.width, im.height, im.rgb.data()});
        mt_scene_desc d{};
        d.meshes = mesh_descs_.data();
        d.mesh_count = static_cast<uint32_t>(mesh_descs_.size());
        d.images = image_descs_.data();
        d.image_count = static_cast<uint32_t>(image_descs_.size());
        d.materials = materials.data();
        d.material_count = static_cast<uint32_t>(materials.size());
        d.instances = instances.data();
        d.instance_count = static_cast<uint32_t>(instances.size());
        d.lights = lights.data();
        d.light_count = static_cast<uint32_t>(lights.size());
        if (!env_rgb8.empty() || !env_rgbf.empty()) {
            d.env.format = env_rgbf.empty() ? MT_ENV_RGB8 : MT_ENV_RGB32F;
            d.env.width = env_width;
            d.env.height = env_height;
            d.env.texels = env_rgbf.empty() ? static_cast<const void*>(env_rgb8.data())
                                            : static_cast<const void*>(env_rgbf.data());
            d.env.filter = MT_FILTER_LINEAR;
        }
        d.bvh = bvh;
        return d;
    }

  private:
    mutable std::vector<mt_mesh_desc> mesh_descs_;
    mutable std::vector<mt_image_desc> image_descs_;
};

namespace detail {

using Mat4 = std::array<float, 16>;  // column-major, element (r, c) at [c * 4 + r]

inline Mat4 Identity() {
    Mat4 m{};
    m[0] = m[5] = m[10] = m[15] = 1.0f;
    return m;
}
// glm's mat4 product: column j of a*b = ((a0*b0j + a1*b1j) + a2*b2j) + a3*b3j
inline Mat4 Mul(const Mat4& a, const Mat4& b) {
    Mat4 r{};
    for (int j = 0; j < 4; ++j)
        for (int i = 0; i < 4; ++i) {
            float s = a[0 * 4 + i] * b[j * 4 + 0];
            s = s + a[1 * 4 + i] * b[j * 4 + 1];
            s = s + a[2 * 4 + i] * b[j * 4 + 2];
            s = s + a[3 * 4 + i] * b[j * 4 + 3];
            r[j * 4 + i] = s;
        }
    return r;
}

class Gltf {
  public:
    explicit Gltf(const std::filesystem::path& path) : base_(path.parent_path()) {
        const std::vector<uint8_t> text = image::ReadFile(path);
        j_ = json::Value::Parse(std::string(text.begin(), text.end()));
    }

    LoadedScene Load(std::size_t scene_index, std::size_t camera_index, bool ext_sky) {
        LoadedScene out;
        std::vector<mt_camera_desc> cameras;
        std::map<std::pair<std::size_t, std::size_t>, uint32_t> mesh_ids;  // dedupe (Scene.cpp:14-22)
        const json::Value& scene = j_["scenes"][scene_index];
        if (scene.Has("nodes"))
            for (std::size_t k = 0; k < scene["nodes"].Size(); ++k)
                Walk(scene["nodes"][k].Index(), Identity(), out, cameras, mesh_ids);
        if (cameras.empty()) throw std::runtime_error("No camera found in the scene");
        out.camera = camera_index < cameras.size() ? cameras[camera_index] : cameras[0];
        out.images = std::move(images_);
        if (ext_sky) {
            const json::Value* ext = j_.Find("extensions");
            const json::Value* sky = ext ? ext->Find("EXT_sky") : nullptr;
            if (sky && sky->Has("sky_texture")) {
                const json::Value& im = j_["images"][j_["textures"][(*sky)["sky_texture"].Index()]["source"].Index()];
                if (!im.Has("uri")) throw std::runtime_error("Only URI-based images are supported");
                out.SetEnvironment(base_ / im["uri"].Str());
            }
        }
        return out;
    }

  private:
    std::filesystem::path base_;
    json::Value j_;
    std::map<std::size_t, std::vector<uint8_t>> buffers_;
    std::map<std::string, int32_t> image_ids_;  // path -> image index (Texture.cpp:17-25)
    std::vector<image::Rgb8> images_;

    const std::vector<uint8_t>& Buffer(std::size_t i) {
        auto it = buffers_.find(i);
        if (it != buffers_.end()) return it->second;
        const json::Value& b = j_["buffers"][i];
        const std::string& uri = b["uri"].Str();
        if (uri.rfind("data:", 0) == 0) throw std::runtime_error("Data URI buffers are not supported");
        std::vector<uint8_t> data = image::ReadFile(base_ / uri);
        const std::size_t len = b["byteLength"].Index();
        if (data.size() < len) throw std::runtime_error("Failed to read buffer file");
        data.resize(len);
        return buffers_.emplace(i, std::move(data)).first->second;
    }
    // accessor i as `comps` tightly packed components of `comp_type` each (GLTF.cpp:41-214)
    template <typename T>
    std::vector<T> Accessor(std::size_t i, const char* type, int comps, int comp_type) {
        static const std::map<std::string, int> kComps = {{"SCALAR", 1}, {"VEC2", 2}, {"VEC3", 3}, {"VEC4", 4}};
        const json::Value& a = j_["accessors"][i];
        const int ct = static_cast<int>(a["componentType"].Num());
        if (!kComps.count(a["type"].Str()) || (ct != 5126 && ct != 5123 && ct != 5125))
            throw std::runtime_error("Unsupported accessor type");
        if (a["type"].Str() != type || ct != comp_type)
            throw std::runtime_error("Attribute type does not match expected type");
        const json::Value& bv = j_["bufferViews"][a["bufferView"].Index()];
        const std::vector<uint8_t>& data = Buffer(bv["buffer"].Index());
        const std::size_t off = static_cast<std::size_t>(a.NumOr("byteOffset", 0) + bv.NumOr("byteOffset", 0));
        const std::size_t n = a["count"].Index() * std::size_t(comps);
        if (off + n * sizeof(T) > data.size()) throw std::runtime_error("Accessor exceeds its buffer");
        std::vector<T> out(n);
        std::memcpy(out.data(), data.data() + off, n * sizeof(T));
        return out;
    }
    std::vector<uint32_t> Indices(std::size_t i) {
        const json::Value& a = j_["accessors"][i];
        const int ct = static_cast<int>(a["componentType"].Num());
        if (a["type"].Str() == "SCALAR" && ct == 5125) return Accessor<uint32_t>(i, "SCALAR", 1, 5125);
        if (a["type"].Str() == "SCALAR" && ct == 5123) {
            const std::vector<uint16_t> s = Accessor<uint16_t>(i, "SCALAR", 1, 5123);
            return std::vector<uint32_t>(s.begin(), s.end());
        }
        throw std::runtime_error("Indices accessor has unsupported component type");
    }
    // the de-indexed soup of mesh mi, primitive pi (GLTF.cpp:308-409, Mesh.cpp:9-35)
    LoadedScene::Mesh MakeMesh(std::size_t mi, std::size_t pi) {
        const json::Value& m = j_["meshes"][mi];
        const json::Value& prim = m["primitives"][pi];
        if (!prim.Has("attributes") || !prim.Has("indices"))
            throw std::runtime_error("Invalid GLTF primitive data: missing attributes or indices");
        if (prim.NumOr("mode", 4) != 4) throw std::runtime_error("Only triangle primitives (mode 4) are supported");
        const json::Value& at = prim["attributes"];
        if (!at.Has("POSITION") || !at.Has("NORMAL")) throw std::runtime_error("POSITION and NORMAL attributes are required");
        const std::vector<uint32_t> idx = Indices(prim["indices"].Index());
        const std::vector<float> pos = Accessor<float>(at["POSITION"].Index(), "VEC3", 3, 5126);
        const std::vector<float> nrm = Accessor<float>(at["NORMAL"].Index(), "VEC3", 3, 5126);
        std::vector<float> uv;
        if (at.Has("TEXCOORD_0")) uv = Accessor<float>(at["TEXCOORD_0"].Index(), "VEC2", 2, 5126);
        if (idx.size() % 3) throw std::runtime_error("Index count is not a multiple of 3");
        LoadedScene::Mesh out;
        out.name = m.Has("name") ? m["name"].Str() : "Unnamed_Mesh";
        for (uint32_t v : idx) {
            if (std::size_t(v) * 3 + 2 >= pos.size() || std::size_t(v) * 3 + 2 >= nrm.size() ||
                (!uv.empty() && std::size_t(v) * 2 + 1 >= uv.size()))
                throw std::runtime_error("Index out of range of the vertex data");
            out.positions.insert(out.positions.end(), &pos[std::size_t(v) * 3], &pos[std::size_t(v) * 3] + 3);
            out.normals.insert(out.normals.end(), &nrm[std::size_t(v) * 3], &nrm[std::size_t(v) * 3] + 3);
            if (!uv.empty()) out.texcoord0.insert(out.texcoord0.end(), &uv[std::size_t(v) * 2], &uv[std::size_t(v) * 2] + 2);
        }
        return out;
    }
    int32_t ImageIndex(std::size_t ii) {
        const json::Value& im = j_["images"][ii];
        if (!im.Has("uri")) throw std::runtime_error("Only URI-based images are supported");
        const std::string path = (base_ / im["uri"].Str()).lexically_normal().string();
        auto it = image_ids_.find(path);
        if (it != image_ids_.end()) return it->second;
        const int32_t id = static_cast<int32_t>(images_.size());
        images_.push_back(image::LoadImageRgb8(path));
        image_ids_.emplace(path, id);
        return id;
    }
    // Texture (GLTF.cpp:216-273): sampler minFilter picks nearest or linear
    mt_texture_desc TextureInfo(const json::Value& info) {
        const json::Value& t = j_["textures"][info["index"].Index()];
        mt_texture_desc d{-1, MT_FILTER_LINEAR, {0.0f, 0.0f}, {1.0f, 1.0f}};
        if (t.Has("sampler")) {
            const int mf = static_cast<int>(j_["samplers"][t["sampler"].Index()].NumOr("minFilter", 9729));
            if (mf == 9728 || mf == 9984 || mf == 9985) d.filter = MT_FILTER_NEAREST;
            else if (mf == 9729 || mf == 9986 || mf == 9987) d.filter = MT_FILTER_LINEAR;
            else throw std::runtime_error("Unsupported filter mode in texture sampler");
        }
        if (info.NumOr("texCoord", 0) != 0) throw std::runtime_error("Only TEXCOORD_0 is supported for texture info");
        d.image = ImageIndex(t["source"].Index());
        const json::Value* ext = info.Find("extensions");
        const json::Value* tt = ext ? ext->Find("KHR_texture_transform") : nullptr;
        if (tt) {
            if (tt->Has("rotation")) throw std::runtime_error("KHR_texture_transform rotation is not supported");
            if (tt->Has("texCoord")) throw std::runtime_error("KHR_texture_transform texCoord override is not supported");
            if (tt->Has("offset")) {
                d.offset[0] = static_cast<float>((*tt)["offset"][0].Num());
                d.offset[1] = static_cast<float>((*tt)["offset"][1].Num());
            }
            if (tt->Has("scale")) {
                d.scale[0] = static_cast<float>((*tt)["scale"][0].Num());
                d.scale[1] = static_cast<float>((*tt)["scale"][1].Num());
            }
        }
        return d;
    }
    static mt_material_desc EmptyMaterial() {  // Loader::Material{} (Types.hpp:65-82)
        mt_material_desc m{};
        const mt_texture_desc none{-1, MT_FILTER_LINEAR, {0.0f, 0.0f}, {1.0f, 1.0f}};
        for (int i = 0; i < 4; ++i) m.base_color_factor[i] = 1.0f;
        m.normal_scale = 1.0f;
        m.metallic_factor = 0.5f;
        m.roughness_factor = 0.5f;
        m.occlusion_strength = 0.0f;
        m.base_color_tex = m.metallic_roughness_tex = m.normal_tex = m.occlusion_tex = m.emissive_tex = none;
        return m;
    }
    // LoadMeshMaterial (GLTF.cpp:411-506)
    mt_material_desc Material(std::size_t mi, std::size_t pi) {
        const json::Value& prim = j_["meshes"][mi]["primitives"][pi];
        if (!prim.Has("material")) return EmptyMaterial();
        const json::Value& md = j_["materials"][prim["material"].Index()];
        if (!md.Has("pbrMetallicRoughness")) return EmptyMaterial();
        const json::Value& pbr = md["pbrMetallicRoughness"];
        mt_material_desc m = EmptyMaterial();
        if (md.Has("emissiveTexture")) m.emissive_tex = TextureInfo(md["emissiveTexture"]);
        float em[3] = {0.0f, 0.0f, 0.0f};
        if (md.Has("emissiveFactor"))
            for (int c = 0; c < 3; ++c) em[c] = static_cast<float>(md["emissiveFactor"][c].Num());
        const json::Value* ext = md.Find("extensions");
        const json::Value* es = ext ? ext->Find("KHR_materials_emissive_strength") : nullptr;
        if (es) {
            const float s = static_cast<float>(es->NumOr("emissiveStrength", 1.0));
            for (float& e : em) e = e * s;
        }
        for (int c = 0; c < 3; ++c) m.emissive_factor[c] = em[c];
        if (pbr.Has("baseColorFactor"))
            for (int c = 0; c < 4; ++c) m.base_color_factor[c] = static_cast<float>(pbr["baseColorFactor"][c].Num());
        m.metallic_factor = static_cast<float>(pbr.NumOr("metallicFactor", 1.0));
        m.roughness_factor = static_cast<float>(pbr.NumOr("roughnessFactor", 1.0));
        if (pbr.Has("baseColorTexture")) m.base_color_tex = TextureInfo(pbr["baseColorTexture"]);
        if (pbr.Has("metallicRoughnessTexture")) m.metallic_roughness_tex = TextureInfo(pbr["metallicRoughnessTexture"]);
        if (md.Has("normalTexture")) m.normal_tex = TextureInfo(md["normalTexture"]);
        m.normal_scale = static_cast<float>(md.NumOr("normalScale", 1.0));  // material-level key (GLTF.cpp:480)
        if (md.Has("occlusionTexture")) m.occlusion_tex = TextureInfo(md["occlusionTexture"]);
        m.occlusion_strength = static_cast<float>(md.NumOr("occlusionStrength", 1.0));
        if (md.Has("alphaMode")) {
            const std::string& am = md["alphaMode"].Str();
            if (am != "OPAQUE" && am != "MASK" && am != "BLEND") throw std::runtime_error("Unsupported alpha mode in material");
        }
        return m;
    }
    // ComputeNodeTransform (GLTF.cpp:275-306): the matrix, or T * R * S
    Mat4 NodeTransform(const json::Value& nd) const {
        Mat4 m = Identity();
        if (nd.Has("matrix")) {
            for (int i = 0; i < 16; ++i) m[i] = static_cast<float>(nd["matrix"][i].Num());
            return m;
        }
        if (nd.Has("translation"))
            for (int r = 0; r < 3; ++r) m[12 + r] = static_cast<float>(nd["translation"][r].Num());
        if (nd.Has("rotation")) {  // glm::mat4_cast of the unit quaternion (x, y, z, w)
            const float x = static_cast<float>(nd["rotation"][0].Num()), y = static_cast<float>(nd["rotation"][1].Num());
            const float z = static_cast<float>(nd["rotation"][2].Num()), w = static_cast<float>(nd["rotation"][3].Num());
            Mat4 r = Identity();
            r[0] = 1.0f - 2.0f * (y * y + z * z);
            r[4] = 2.0f * (x * y - w * z);
            r[8] = 2.0f * (x * z + w * y);
            r[1] = 2.0f * (x * y + w * z);
            r[5] = 1.0f - 2.0f * (x * x + z * z);
            r[9] = 2.0f * (y * z - w * x);
            r[2] = 2.0f * (x * z - w * y);
            r[6] = 2.0f * (y * z + w * x);
            r[10] = 1.0f - 2.0f * (x * x + y * y);
            m = Mul(m, r);
        }
        if (nd.Has("scale"))
            for (int c = 0; c < 3; ++c) {
                const float s = static_cast<float>(nd["scale"][c].Num());
                for (int r = 0; r < 4; ++r) m[c * 4 + r] = m[c * 4 + r] * s;
            }
        return m;
    }
    // a light at `world`: position from the translation, direction = normalised -z axis
    static mt_light_desc MakeLight(int32_t type, const Mat4& world, const float col[3], float size,
                                   float inner, float outer) {
        mt_light_desc l{};
        l.type = type;
        for (int r = 0; r < 3; ++r) l.position[r] = world[12 + r];
        for (int c = 0; c < 3; ++c) l.color[c] = col[c];
        l.point_size = size;
        l.spot_inner_cos = std::cos(inner);
        l.spot_outer_cos = std::cos(outer);
        const double dx = -double(world[8]), dy = -double(world[9]), dz = -double(world[10]);
        const double n = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (n > 0.0) {
            l.direction[0] = static_cast<float>(dx / n);
            l.direction[1] = static_cast<float>(dy / n);
            l.direction[2] = static_cast<float>(dz / n);
        } else {
            l.direction[0] = 0.0f;
            l.direction[1] = 0.0f;
            l.direction[2] = -1.0f;
        }
        return l;
    }
    static void Color(const json::Value& ld, float col[3]) {
        const float in = static_cast<float>(ld.NumOr("intensity", 1.0));
        for (int c = 0; c < 3; ++c)
            col[c] = (ld.Has("color") ? static_cast<float>(ld["color"][c].Num()) : 1.0f) * in;
    }
    // LoadPointLight / LoadSpotLight / LoadDirectionalLight (GLTF.cpp:508-600)
    mt_light_desc Light(std::size_t li, const Mat4& world) const {
        const json::Value& ld = j_["extensions"]["KHR_lights_punctual"]["lights"][li];
        const std::string type = ld.Has("type") ? ld["type"].Str() : "";
        float col[3];
        Color(ld, col);
        if (type == "point") return MakeLight(MT_LIGHT_POINT, world, col, static_cast<float>(ld.NumOr("size", 0.1)), 0.0f, 0.0f);
        if (type == "directional") return MakeLight(MT_LIGHT_DIRECTIONAL, world, col, 0.1f, 0.0f, 0.0f);
        if (type == "spot") {
            const json::Value* spot = ld.Find("spot");
            const float inner = spot ? static_cast<float>(spot->NumOr("innerConeAngle", 0.0)) : 0.0f;
            const float outer = spot ? static_cast<float>(spot->NumOr("outerConeAngle", 0.7853981633974483)) : 0.7853981633974483f;
            return MakeLight(MT_LIGHT_SPOT, world, col, static_cast<float>(ld.NumOr("size", 0.0)), inner, outer);
        }
        throw std::runtime_error("Unsupported light type in GLTF");
    }
    // LoadAreaLight (GLTF.cpp:602-640)
    mt_light_desc AreaLight(std::size_t ai, const Mat4& world) const {
        const json::Value& ld = j_["extensions"]["EXT_light_attributes"]["areaLights"][ai];
        if (!ld.Has("shape")) throw std::runtime_error("Area light missing 'shape' property");
        if (ld["shape"].Str() != "rectangle" && ld["shape"].Str() != "disc") throw std::runtime_error("Unsupported area light shape");
        float col[3];
        Color(ld, col);
        return MakeLight(MT_LIGHT_AREA, world, col, 0.1f, 0.0f, 0.0f);
    }
    // LoadCamera (GLTF.cpp:642-679)
    mt_camera_desc Camera(std::size_t ci, const Mat4& world) const {
        const json::Value& cd = j_["cameras"][ci];
        mt_camera_desc c{};
        std::copy(world.begin(), world.end(), c.camera_to_world);
        if (cd["type"].Str() == "perspective") {
            const json::Value& p = cd["perspective"];
            c.type = MT_CAMERA_PERSPECTIVE;
            c.yfov = static_cast<float>(p["yfov"].Num());
            c.aspect = static_cast<float>(p.NumOr("aspectRatio", 1.0));
            return c;
        }
        if (cd["type"].Str() == "orthographic") {
            const json::Value& o = cd["orthographic"];
            c.type = MT_CAMERA_ORTHOGONAL;
            c.ortho_size[0] = static_cast<float>(o["xmag"].Num());
            c.ortho_size[1] = static_cast<float>(o["ymag"].Num());
            c.aspect = c.ortho_size[0] / c.ortho_size[1];
            return c;
        }
        throw std::runtime_error("Unsupported camera type in GLTF");
    }
    // depth-first over the node tree: meshes, lights and cameras in node order (the reference
    // collects each kind in its own walk, which visits the nodes in the same order)
    void Walk(std::size_t ni, const Mat4& parent, LoadedScene& out, std::vector<mt_camera_desc>& cameras,
              std::map<std::pair<std::size_t, std::size_t>, uint32_t>& mesh_ids) {
        const json::Value& nd = j_["nodes"][ni];
        const Mat4 world = Mul(parent, NodeTransform(nd));
        if (nd.Has("mesh")) {
            const std::size_t mi = nd["mesh"].Index();
            for (std::size_t pi = 0; pi < j_["meshes"][mi]["primitives"].Size(); ++pi) {
                auto key = std::make_pair(mi, pi);
                auto it = mesh_ids.find(key);
                if (it == mesh_ids.end()) {
                    out.meshes.push_back(MakeMesh(mi, pi));
                    it = mesh_ids.emplace(key, static_cast<uint32_t>(out.meshes.size() - 1)).first;
                }
                out.materials.push_back(Material(mi, pi));
                mt_instance_desc inst{};
                inst.mesh = it->second;
                inst.material = static_cast<uint32_t>(out.materials.size() - 1);
                std::copy(world.begin(), world.end(), inst.transform);
                out.instances.push_back(inst);
            }
        }
        if (const json::Value* ext = nd.Find("extensions")) {
            if (const json::Value* kl = ext->Find("KHR_lights_punctual")) out.lights.push_back(Light((*kl)["light"].Index(), world));
            if (const json::Value* ea = ext->Find("EXT_light_attributes"))
                if (ea->Has("areaLight")) out.lights.push_back(AreaLight((*ea)["areaLight"].Index(), world));
        }
        if (nd.Has("camera")) cameras.push_back(Camera(nd["camera"].Index(), world));
        if (nd.Has("children"))
            for (std::size_t k = 0; k < nd["children"].Size(); ++k)
                Walk(nd["children"][k].Index(), world, out, cameras, mesh_ids);
    }
};

}  // namespace detail

/* Loader::GLTF(path).LoadScene(sceneIndex) + Scene::Scene + the camera of cameraIndex (the
 * first camera when the index is out of range).  extSky: also keep the EXT_sky texture as the
 * environment (GLTF.cpp:902-912), which Scene::Scene drops (Scene.cpp:53-59). */
inline LoadedScene LoadGltf(const std::filesystem::path& path, std::size_t sceneIndex = 0,
                            std::size_t cameraIndex = 0, bool extSky = false) {
    return detail::Gltf(path).Load(sceneIndex, cameraIndex, extSky);
}

/* Config::ConfigOptions (src/CLI/Config.hpp:43-80): the fields the render path reads, with the
 * reference's defaults (the firefly flag is inverted on the command line, Config.cpp:166-170,
 * so its effective default is off), plus the GPU's seed and traversal */
struct CliOptions {
    std::string inputFilename, outputFilename;
    int debugMode = MT_DEBUG_NONE;
    unsigned imageWidth = 1280, imageHeight = 720, samples = 64, bounces = 5;
    unsigned cameraIndex = 0, sceneIndex = 0;
    float evExposure = 8.0f;
    std::string hdriOverride;
    float hdriRotation = 0.0f, hdriSecondaryIntensity = 350.0f, hdriPrimaryIntensity = 350.0f;
    float emissionBaseIntensity = 1.0f;
    bool bvhDisable = false;
    unsigned bvhMaxTriangles = 32, bvhMaxDepth = 32, imageBlockSize = 32;
    bool useStatisticFireflyElimination = false;
    float fireflyEliminationThreshold = 6.0f, terminateEnergy = 0.01f;
    bool quiet = false;
    uint64_t seed = 0;
    int traversal = MT_TRAVERSAL_REFERENCE;
};

/* the options as mt_settings (the GPU additions at their defaults: reference estimator,
 * tone-mapped output, the rotation of the environment ignored as the reference does) */
inline mt_settings ToSettings(const CliOptions& o) {
    mt_settings s{};
    s.width = o.imageWidth;
    s.height = o.imageHeight;
    s.samples = o.samples;
    s.bounces = o.bounces;
    s.ev_exposure = o.evExposure;
    s.hdri_rotation = o.hdriRotation;
    s.hdri_secondary_intensity = o.hdriSecondaryIntensity;
    s.hdri_primary_intensity = o.hdriPrimaryIntensity;
    s.emission_base_intensity = o.emissionBaseIntensity;
    s.firefly_enable = o.useStatisticFireflyElimination ? 1 : 0;
    s.firefly_threshold = o.fireflyEliminationThreshold;
    s.terminate_energy = o.terminateEnergy;
    s.image_block_size = o.imageBlockSize;
    s.seed = o.seed;
    s.mode = MT_MODE_PARITY;
    s.rr_depth = 3;
    s.output = MT_OUTPUT_TONEMAPPED;
    s.debug_mode = o.debugMode;
    s.traversal = o.traversal;
    return s;
}

/* the BVH options of the configuration (Mesh::Mesh reads them from the Config singleton,
 * Mesh.cpp:37-40) */
inline mt_bvh_options ToBvhOptions(const CliOptions& o) {
    mt_bvh_options b{o.bvhDisable ? 1 : 0, o.bvhMaxTriangles, o.bvhMaxDepth,
                     o.traversal == MT_TRAVERSAL_FAST ? 1 : 0, 65536, 0};
    return b;
}

}  // namespace mitrace_gpu

#endif  // MITRACE_GPU_LOADER_HPP
