// pfxh-render: a C++ caller of libpfxhip that renders one still the way
// `pfx_still::render_kept` does (crates/still/src/trace.rs:425-575): build
// the trace, take a fixed number of samples or converge adaptively, denoise
// if asked, and write the film (colour, albedo, normal, denoised) as PFM.
//
//   pfxh-render SCENE.pfxscn --width W --height H [--spp N | --adaptive THRESHOLD,MIN,MAX]
//               [--bounces B] [--seed S] [--denoise none|repo|atrous|learned]
//               [--weights FILE] [--out PREFIX]
//
// SCENE is a PFXSCN02 file (pfx_amd.scene.save_scene). The repo denoiser
// reads OIDN's TZA weights and the learned one a PFXLNET1 file; both default
// to the files beside the library (weights/).
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/pfxhip.h"

namespace {

std::vector<uint8_t> read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot read " + path);
    return std::vector<uint8_t>(std::istreambuf_iterator<char>(f), {});
}

struct SceneFile {
    std::vector<uint8_t> bytes;
    pfxh_scene scene{};
};

// PFXSCN02: magic, 9 u32 counts, camera (12 f32), sun (7 f32), the arrays
// as the pfxh_* structs lay them out, the sky texels.
SceneFile load_scene(const std::string& path) {
    SceneFile s;
    s.bytes = read_file(path);
    const uint8_t* p = s.bytes.data();
    size_t len = s.bytes.size(), at = 120;
    if (len < 120 || memcmp(p, "PFXSCN02", 8) != 0) throw std::runtime_error(path + ": not a PFXSCN02 scene");
    uint32_t n[9];
    memcpy(n, p + 8, sizeof n);
    float cam[12], sun[7];
    memcpy(cam, p + 44, sizeof cam);
    memcpy(sun, p + 92, sizeof sun);
    auto take = [&](size_t count, size_t size) -> const void* {
        if (count > (len - at) / (size ? size : 1)) throw std::runtime_error(path + ": truncated");
        const void* q = count ? p + at : nullptr;
        at += count * size;
        return q;
    };
    pfxh_scene& sc = s.scene;
    sc.triangles = (const pfxh_triangle*)take(n[0], sizeof(pfxh_triangle));
    sc.triangle_count = n[0];
    sc.world_details = (const pfxh_triangle_detail*)take(n[1], sizeof(pfxh_triangle_detail));
    sc.world_detail_count = n[1];
    sc.mesh_triangles = (const pfxh_triangle_detail*)take(n[2], sizeof(pfxh_triangle_detail));
    sc.mesh_triangle_count = n[2];
    sc.meshes = (const pfxh_mesh*)take(n[3], sizeof(pfxh_mesh));
    sc.mesh_count = n[3];
    sc.instances = (const pfxh_instance*)take(n[4], sizeof(pfxh_instance));
    sc.instance_count = n[4];
    sc.materials = (const pfxh_material*)take(n[5], sizeof(pfxh_material));
    sc.material_count = n[5];
    sc.lights = (const pfxh_light*)take(n[6], sizeof(pfxh_light));
    sc.light_count = n[6];
    sc.sky.width = n[7];
    sc.sky.height = n[8];
    sc.sky.texels = (const float*)take((size_t)n[7] * n[8], 16);
    memcpy(sc.camera.origin, cam, 12);
    memcpy(sc.camera.forward, cam + 3, 12);
    memcpy(sc.camera.right, cam + 6, 12);
    memcpy(sc.camera.up, cam + 9, 12);
    memcpy(sc.sun.direction, sun, 12);
    memcpy(sc.sun.color, sun + 3, 12);
    sc.sun.intensity = sun[6];
    return s;
}

// PFM, RGB, little-endian, rows bottom to top
void write_pfm(const std::string& path, const std::vector<float>& rgba, uint32_t w, uint32_t h) {
    FILE* f = fopen(path.c_str(), "wb");
    if (!f) throw std::runtime_error("cannot write " + path);
    fprintf(f, "PF\n%u %u\n-1.0\n", w, h);
    std::vector<float> row((size_t)w * 3);
    for (uint32_t y = h; y-- > 0;) {
        for (uint32_t x = 0; x < w; ++x)
            for (int c = 0; c < 3; ++c) row[(size_t)x * 3 + c] = rgba[((size_t)y * w + x) * 4 + c];
        fwrite(row.data(), 4, row.size(), f);
    }
    fclose(f);
}

void check(pfxh_status st, const char* what, const char* err) {
    if (st != PFXH_OK) throw std::runtime_error(std::string(what) + ": " + err + " (status " + std::to_string(st) + ")");
}

std::string beside_library(const char* argv0, const std::string& rel) {
    std::string self(argv0);
    size_t slash = self.rfind('/');
    return (slash == std::string::npos ? std::string(".") : self.substr(0, slash)) + "/" + rel;
}

double ms_since(std::chrono::steady_clock::time_point t) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t).count();
}

}  // namespace

int main(int argc, char** argv) {
    try {
        if (argc < 2) {
            fprintf(stderr, "usage: %s SCENE.pfxscn --width W --height H [--spp N | --adaptive T,MIN,MAX] "
                            "[--bounces B] [--seed S] [--denoise none|repo|atrous|learned] [--weights FILE] "
                            "[--out PREFIX]\n", argv[0]);
            return 2;
        }
        std::string scene_path = argv[1], denoise = "none", weights, out = "still";
        uint32_t width = 512, height = 512, spp = 16, bounces = 6, seed = 1;
        bool adaptive = false;
        pfxh_adaptive a{0.01f, 16, 64, 1.5f};
        for (int i = 2; i < argc; ++i) {
            std::string k = argv[i];
            auto next = [&]() -> std::string {
                if (i + 1 >= argc) throw std::runtime_error(k + " needs a value");
                return argv[++i];
            };
            if (k == "--width") width = (uint32_t)std::stoul(next());
            else if (k == "--height") height = (uint32_t)std::stoul(next());
            else if (k == "--spp") spp = (uint32_t)std::stoul(next());
            else if (k == "--bounces") bounces = (uint32_t)std::stoul(next());
            else if (k == "--seed") seed = (uint32_t)std::stoul(next());
            else if (k == "--denoise") denoise = next();
            else if (k == "--weights") weights = next();
            else if (k == "--out") out = next();
            else if (k == "--adaptive") {
                adaptive = true;
                std::string v = next();
                if (sscanf(v.c_str(), "%f,%u,%u", &a.threshold, &a.min_samples, &a.max_samples) != 3)
                    throw std::runtime_error("--adaptive wants THRESHOLD,MIN,MAX");
            } else throw std::runtime_error("unknown option " + k);
        }
        if (spp == 0 && !adaptive) throw std::runtime_error("a render needs at least one sample");
        SceneFile sf = load_scene(scene_path);
        pfxh_detail detail{};
        detail.bounces = pfxh_bounces{bounces, bounces, bounces, bounces};
        char err[1024] = {0};
        auto t0 = std::chrono::steady_clock::now();
        pfxh_trace* trace = nullptr;
        check(pfxh_trace_new(&sf.scene, &detail, width, height, &trace, err, sizeof err), "build", err);
        printf("built %ux%u in %.1f ms (%.2f GiB on the GPU)\n", width, height, ms_since(t0),
               pfxh_trace_gpu_bytes(trace) / 1073741824.0);
        t0 = std::chrono::steady_clock::now();
        if (adaptive) {
            pfxh_converged conv{};
            check(pfxh_trace_sample_adaptive(trace, &a, seed, &conv, err, sizeof err), "sample", err);
        } else {
            check(pfxh_trace_sample(trace, spp, seed, err, sizeof err), "sample", err);
        }
        double traced = ms_since(t0);
        pfxh_stats st{};
        pfxh_trace_stats(trace, &st);
        printf("traced %u spp in %.1f ms: %.1f Mrays/s\n", pfxh_trace_samples(trace), traced,
               (double)(st.closest_rays + st.shadow_rays) / (traced * 1e3));
        size_t n = (size_t)width * height * 4;
        std::vector<float> color(n), albedo(n), normal(n), den;
        check(pfxh_trace_readback(trace, color.data(), albedo.data(), normal.data(), err, sizeof err), "readback",
              err);
        write_pfm(out + ".color.pfm", color, width, height);
        write_pfm(out + ".albedo.pfm", albedo, width, height);
        write_pfm(out + ".normal.pfm", normal, width, height);
        if (denoise != "none") {
            den.resize(n);
            t0 = std::chrono::steady_clock::now();
            if (denoise == "atrous") {
                check(pfxh_trace_atrous(trace, den.data(), err, sizeof err), "denoise", err);
            } else if (denoise == "repo") {
                std::vector<uint8_t> w = read_file(weights.empty() ? beside_library(argv[0], "weights/rt_hdr_alb_nrm.tza")
                                                                   : weights);
                pfxh_denoiser* d = nullptr;
                check(pfxh_denoiser_new(w.data(), w.size(), &d, err, sizeof err), "denoiser", err);
                check(pfxh_trace_denoise(trace, d, den.data(), err, sizeof err), "denoise", err);
                pfxh_denoiser_free(d);
            } else if (denoise == "learned") {
                std::vector<uint8_t> w = read_file(weights.empty() ? beside_library(argv[0], "weights/learned_terrain.pfxl")
                                                                   : weights);
                if (w.size() < 16 || memcmp(w.data(), "PFXLNET1", 8) != 0) throw std::runtime_error("not a PFXLNET1 file");
                uint64_t count = 0;
                memcpy(&count, w.data() + 8, 8);
                if (w.size() != 16 + count * 4) throw std::runtime_error("truncated PFXLNET1 file");
                pfxh_learner* l = nullptr;
                check(pfxh_learner_new(0, &l, err, sizeof err), "learner", err);
                check(pfxh_learner_set_params(l, (const float*)(w.data() + 16), count, err, sizeof err), "weights",
                      err);
                check(pfxh_trace_learned_denoise(trace, l, den.data(), err, sizeof err), "denoise", err);
                pfxh_learner_free(l);
            } else {
                throw std::runtime_error("unknown denoiser " + denoise);
            }
            printf("denoised (%s) in %.1f ms\n", denoise.c_str(), ms_since(t0));
            write_pfm(out + ".denoised.pfm", den, width, height);
        }
        pfxh_trace_free(trace);
        return 0;
    } catch (const std::exception& e) {
        fprintf(stderr, "pfxh-render: %s\n", e.what());
        return 1;
    }
}
