// Drives LeniaHipModule (life-simulation_amd/binding) the way SceneRunner::step
// drives a module (LifeCore/Sim/SceneRunner.cpp:146-215): configure, setup,
// reset with the scene seed, then one encode per frame.  Built by
// tests/test_binding.py with g++ against include/lenia_mi355x.h.
//
//   binding_driver describe <params.json> <name> <W> <H>
//       prints what setup() resolves (no GPU call)
//   binding_driver run <params.json> <name> <W> <H> <seed> <frames> <substeps>
//                      <bus.txt> <state.bin> <rgba.bin>
//       steps the module and writes its field (float32) and colour layer
//       (RGBA16F); bus.txt holds "module.key value" lines, the values the
//       ParameterBus returns every frame
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include <hip/hip_runtime_api.h>

#include "lenia_hip_module.h"

namespace {

struct MapBus : life_hip::ParameterBus {
    std::map<std::string, float> v;
    float value(const std::string &key, float fallback) const override {
        auto it = v.find(key);
        return it == v.end() ? fallback : it->second;
    }
};

std::string slurp(const char *path) {
    std::ifstream f(path, std::ios::binary);
    std::stringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

// top-level keys of a JSON object (enough for the test's params files:
// a key is a string directly inside the outermost braces followed by ':')
std::set<std::string> top_keys(const std::string &s) {
    std::set<std::string> keys;
    int depth = 0;
    for (size_t i = 0; i < s.size(); ++i) {
        char c = s[i];
        if (c == '"') {
            size_t j = i + 1;
            std::string str;
            while (j < s.size() && s[j] != '"') {
                if (s[j] == '\\' && j + 1 < s.size()) ++j;
                str += s[j++];
            }
            size_t k = j + 1;
            while (k < s.size() && (s[k] == ' ' || s[k] == '\n' || s[k] == '\t' || s[k] == '\r')) ++k;
            if (depth == 1 && k < s.size() && s[k] == ':') keys.insert(str);
            i = j;
        } else if (c == '{' || c == '[') {
            ++depth;
        } else if (c == '}' || c == ']') {
            --depth;
        }
    }
    return keys;
}

bool write_file(const char *path, const void *data, size_t bytes) {
    FILE *f = std::fopen(path, "wb");
    if (!f) return false;
    const bool ok = std::fwrite(data, 1, bytes, f) == bytes;
    return std::fclose(f) == 0 && ok;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 6) {
        std::fprintf(stderr, "usage: %s describe|run params.json name W H ...\n", argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    life_hip::ModuleParams params;
    params.text = slurp(argv[2]);
    params.keys = top_keys(params.text);
    const std::string name = argv[3];
    const uint32_t W = (uint32_t)std::atoi(argv[4]), H = (uint32_t)std::atoi(argv[5]);
    if (mode == "describe") {
        lenia_world_desc d;
        lenia_colormap cm;
        if (LeniaHipModule::resolve(name, params.text, W, H, &d, &cm)) {
            std::printf("error %s\n", lenia_last_error());
            return 1;
        }
        std::printf("width %u\nheight %u\nconvMode %u\nnKernels %u\nradius %u\ndt %.9g\n", d.width, d.height,
                    d.convMode, d.nKernels, d.radius, (double)d.dt);
        std::printf("growthMode %u\ninjectAmount %.9g\ninjectCount %u\ninitMode %u\n", d.growthMode,
                    (double)d.injectAmount, d.injectCount, d.initMode);
        for (uint32_t k = 0; k < d.nKernels; ++k)
            std::printf("kernel %u %.9g %.9g %.9g %.9g %u\n", k, (double)d.kernels[k].radius,
                        (double)d.kernels[k].mu, (double)d.kernels[k].sigma, (double)d.kernels[k].weight,
                        d.kernels[k].nBetas);
        std::printf("palette_d %.9g %.9g %.9g\n", (double)cm.d[0], (double)cm.d[1], (double)cm.d[2]);
        return 0;
    }
    if (mode != "run" || argc < 12) {
        std::fprintf(stderr, "run needs: seed frames substeps bus.txt state.bin rgba.bin\n");
        return 2;
    }
    const uint32_t seed = (uint32_t)std::strtoul(argv[6], nullptr, 10);
    const uint32_t frames = (uint32_t)std::atoi(argv[7]);
    const uint32_t substeps = (uint32_t)std::atoi(argv[8]);
    MapBus bus;
    {
        std::ifstream f(argv[9]);
        std::string key;
        float v;
        while (f >> key >> v) bus.v[key] = v;
    }
    life_hip::SimulationContext ctx;
    ctx.params = &bus;
    ctx.width = W;
    ctx.height = H;
    ctx.substeps = substeps;
    ctx.seed = seed;
    LeniaHipModule m;
    m.configure(name, params);
    m.setup(ctx);
    if (!m.ok()) return 1;
    m.reset(seed);  // SceneRunner: first appearance at frame 0, seed ^ 0 * 2654435761
    for (uint32_t f = 0; f < frames; ++f) {
        ctx.frameIndex = f;
        m.encode(ctx);
        if (!m.ok()) return 1;
    }
    if (lenia_ctx_sync(m.context())) return 1;
    const life_hip::DeviceImage field = m.outputField(), tex = m.outputTexture();
    std::vector<float> state((size_t)field.width * field.height);
    std::vector<uint16_t> rgba((size_t)tex.width * tex.height * 4);
    if (hipMemcpy(state.data(), field.data, state.size() * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(rgba.data(), tex.data, rgba.size() * sizeof(uint16_t), hipMemcpyDeviceToHost) != hipSuccess)
        return 1;
    if (!write_file(argv[10], state.data(), state.size() * sizeof(float)) ||
        !write_file(argv[11], rgba.data(), rgba.size() * sizeof(uint16_t)))
        return 1;
    std::printf("frames %u state %ux%u layer %ux%u\n", frames, field.width, field.height, tex.width, tex.height);
    return 0;
}
