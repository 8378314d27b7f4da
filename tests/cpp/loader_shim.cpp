// Test driver of the C++ scene loader (include/mitrace_gpu_loader.hpp): exposes LoadGltf,
// ToDesc and the image readers over a C interface for tests/test_cpp_loader.py.  Needs no GPU
// and no library: the loader is header-only.
#include <cstring>
#include <exception>
#include <string>

#include "../../include/mitrace_gpu_loader.hpp"

namespace {
void copy_message(const char* msg, char* err, int errlen) {
    if (!err || errlen <= 0) return;
    std::strncpy(err, msg, static_cast<std::size_t>(errlen) - 1);
    err[errlen - 1] = '\0';
}
}  // namespace

extern "C" {

// loads a glTF scene (and an environment override when env_path is not empty); returns the
// loaded scene (free with ldr_close) or null with the exception's message in err
void* ldr_open(const char* path, int scene, int camera, int ext_sky, const char* env_path, char* err, int errlen) {
    try {
        auto* s = new mitrace_gpu::LoadedScene(mitrace_gpu::LoadGltf(path, static_cast<std::size_t>(scene),
                                                                     static_cast<std::size_t>(camera), ext_sky != 0));
        if (env_path && *env_path) s->SetEnvironment(env_path);
        return s;
    } catch (const std::exception& e) {
        copy_message(e.what(), err, errlen);
        return nullptr;
    }
}

void ldr_close(void* s) { delete static_cast<mitrace_gpu::LoadedScene*>(s); }

// the scene's C-ABI description (pointing into the loaded scene) and its camera
void ldr_desc(void* s, mt_scene_desc* desc, mt_camera_desc* camera) {
    auto* ls = static_cast<mitrace_gpu::LoadedScene*>(s);
    *desc = ls->ToDesc();
    *camera = ls->camera;
}

// an image file as the loader reads it: kind 0 = 8-bit RGB (PNG), 1 = Radiance .hdr floats.
// With out == null only the size is returned.
int ldr_image(const char* path, int kind, uint32_t* w, uint32_t* h, void* out, char* err, int errlen) {
    try {
        if (kind == 0) {
            const auto im = mitrace_gpu::image::LoadImageRgb8(path);
            *w = im.width;
            *h = im.height;
            if (out) std::memcpy(out, im.rgb.data(), im.rgb.size());
        } else {
            const auto im = mitrace_gpu::image::LoadHdr(path);
            *w = im.width;
            *h = im.height;
            if (out) std::memcpy(out, im.rgb.data(), im.rgb.size() * sizeof(float));
        }
        return 0;
    } catch (const std::exception& e) {
        copy_message(e.what(), err, errlen);
        return 1;
    }
}

// the CLI's options -> settings (ToSettings) for argument defaults checks
void ldr_default_settings(mt_settings* out) { *out = mitrace_gpu::ToSettings(mitrace_gpu::CliOptions{}); }

}  // extern "C"
