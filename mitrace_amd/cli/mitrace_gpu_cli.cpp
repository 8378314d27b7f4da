// mitrace_gpu_cli -- the reference's command line (src/main.cpp, src/CLI/Config.cpp) on the GPU
// Tracer: load a glTF scene with the C++ loader, render it with mitrace_gpu::Tracer while a
// stat thread prints progress, and save the image by the output file's extension
// (.png / .jpg / .hdr).
//
//   mitrace_gpu_cli scene.gltf -o out.png -w 1920 -h 1080 -s 256 -b 5 [options]
//
// The options keep the reference's names and defaults (Config.cpp:45-214); --seed and --fast
// are additions (the counter RNG key, and the SAH traversal of MT_TRAVERSAL_FAST).  As in the
// reference, --disable-firefly-elimination is stored into useStatisticFireflyElimination as a
// plain flag, so it turns the elimination on (Config.cpp:166-170).
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <string>
#include <thread>

#include "mitrace_gpu_loader.hpp"
#include "mitrace_gpu_tracer.hpp"

namespace {

[[noreturn]] void Usage(const char* msg) {
    std::fprintf(stderr, "mitrace_gpu_cli: %s\n", msg);
    std::fprintf(stderr,
                 "usage: mitrace_gpu_cli input.gltf [-o out.png|.jpg|.hdr] [-w W] [-h H] [-s SPP] [-b BOUNCES]\n"
                 "  [-c CAMERA] [--scene N] [-e EV] [-hdri FILE] [--hdri-rotation DEG]\n"
                 "  [--hdri-secondary-intensity X] [--hdri-primary-intensity X] [--emission-base-intensity X]\n"
                 "  [--disable-bvh] [--bvh-max-triangles N] [--bvh-max-depth N] [--image-block-size N]\n"
                 "  [--disable-firefly-elimination] [--firefly-threshold X] [--terminate-energy X]\n"
                 "  [--seed N] [--fast] [-q]\n");
    std::exit(EXIT_FAILURE);
}

mitrace_gpu::CliOptions Parse(int argc, char** argv) {
    mitrace_gpu::CliOptions o;
    auto value = [&](int& i) -> std::string {
        if (i + 1 >= argc) Usage((std::string("missing value for ") + argv[i]).c_str());
        return argv[++i];
    };
    auto num = [&](int& i) { return std::stod(value(i)); };
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        if (a == "-o" || a == "--output") o.outputFilename = value(i);
        else if (a == "-w" || a == "--width") o.imageWidth = static_cast<unsigned>(num(i));
        else if (a == "-h" || a == "--height") o.imageHeight = static_cast<unsigned>(num(i));
        else if (a == "-s" || a == "--samples") o.samples = static_cast<unsigned>(num(i));
        else if (a == "-b" || a == "--bounces") o.bounces = static_cast<unsigned>(num(i));
        else if (a == "-c" || a == "--camera") o.cameraIndex = static_cast<unsigned>(num(i));
        else if (a == "--scene") o.sceneIndex = static_cast<unsigned>(num(i));
        else if (a == "-e" || a == "--ev-exposure") o.evExposure = static_cast<float>(num(i));
        else if (a == "-hdri" || a == "--hdri-override") o.hdriOverride = value(i);
        else if (a == "--hdri-rotation") o.hdriRotation = static_cast<float>(num(i));
        else if (a == "--hdri-secondary-intensity") o.hdriSecondaryIntensity = static_cast<float>(num(i));
        else if (a == "--hdri-primary-intensity") o.hdriPrimaryIntensity = static_cast<float>(num(i));
        else if (a == "--emission-base-intensity") o.emissionBaseIntensity = static_cast<float>(num(i));
        else if (a == "--disable-bvh") o.bvhDisable = true;
        else if (a == "--bvh-max-triangles") o.bvhMaxTriangles = static_cast<unsigned>(num(i));
        else if (a == "--bvh-max-depth") o.bvhMaxDepth = static_cast<unsigned>(num(i));
        else if (a == "--image-block-size") o.imageBlockSize = static_cast<unsigned>(num(i));
        else if (a == "--disable-firefly-elimination") o.useStatisticFireflyElimination = true;
        else if (a == "--firefly-threshold") o.fireflyEliminationThreshold = static_cast<float>(num(i));
        else if (a == "--terminate-energy") o.terminateEnergy = static_cast<float>(num(i));
        else if (a == "--seed") o.seed = std::stoull(value(i));
        else if (a == "--fast") o.traversal = MT_TRAVERSAL_FAST;
        else if (a == "-q" || a == "--quiet") o.quiet = true;
        else if (!a.empty() && a[0] == '-') Usage(("unknown option " + a).c_str());
        else if (o.inputFilename.empty()) o.inputFilename = a;
        else Usage("more than one input file");
    }
    if (o.inputFilename.empty()) Usage("no input file");
    if (o.outputFilename.empty()) o.outputFilename = "render.png";
    return o;
}

}  // namespace

int main(int argc, char** argv) {
    const mitrace_gpu::CliOptions cfg = Parse(argc, argv);
    const std::string& out = cfg.outputFilename;
    const auto dot = out.rfind('.');
    const std::string filetype = dot == std::string::npos ? "" : out.substr(dot);
    if (filetype != ".png" && filetype != ".jpg" && filetype != ".hdr") {  // main.cpp:90-95
        std::fprintf(stderr, "Invalid output file type\n");
        return EXIT_FAILURE;
    }
    try {
        using clock = std::chrono::steady_clock;
        const auto t0 = clock::now();
        mitrace_gpu::LoadedScene loaded = mitrace_gpu::LoadGltf(cfg.inputFilename, cfg.sceneIndex, cfg.cameraIndex);
        if (!cfg.hdriOverride.empty()) loaded.SetEnvironment(cfg.hdriOverride);
        loaded.bvh = mitrace_gpu::ToBvhOptions(cfg);
        const mitrace_gpu::Scene scene(loaded.ToDesc(), loaded.camera);
        const double load_s = std::chrono::duration<double>(clock::now() - t0).count();
        if (!cfg.quiet) std::printf("Loading took %.2f seconds\n", load_s);

        const mt_settings st = mitrace_gpu::ToSettings(cfg);
        auto buf = std::make_shared<mitrace_gpu::RenderBuffer>(st.width, st.height);
        mitrace_gpu::Tracer tracer(buf, scene, st);
        std::atomic<bool> stop{false};
        // StatThread (main.cpp:41-77): progress once a second until the render is done
        std::thread stats([&]() {
            while (!stop.load()) {
                for (int k = 0; k < 10 && !stop.load(); ++k) std::this_thread::sleep_for(std::chrono::milliseconds(100));
                if (stop.load() || cfg.quiet) continue;
                const auto s = tracer.GetStats();
                std::printf("Progress: %5.1f%%, elapsed %.1f s, remaining %.1f s\n", 100.0f * s.progress,
                            s.timeElapsed, s.estimatedTimeRemaining);
                std::fflush(stdout);
            }
        });
        const auto r0 = clock::now();
        tracer.StartRender();
        tracer.WaitForRender();
        stop.store(true);
        stats.join();
        const double secs = std::chrono::duration<double>(clock::now() - r0).count();
        const auto s = tracer.GetStats();
        if (!cfg.quiet)
            std::printf("Render completed in %.2f seconds, rays: %.2fM (%.2fM/s), samples: %.2fM (%.2fM/s)\n", secs,
                        s.rays / 1e6, s.rays / 1e6 / secs, s.samples / 1e6, s.samples / 1e6 / secs);
        if (filetype == ".jpg") buf->SaveToFileJPG(out);
        else if (filetype == ".png") buf->SaveToFilePNG(out);
        else buf->SaveToFileHDR(out);
    } catch (const std::exception& e) {  // main.cpp:29-30: runtime errors end the program
        std::fprintf(stderr, "Error: %s\n", e.what());
        return EXIT_FAILURE;
    }
    return EXIT_SUCCESS;
}
