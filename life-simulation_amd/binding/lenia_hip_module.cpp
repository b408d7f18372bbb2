// LeniaHipModule over the C ABI; see lenia_hip_module.h.
#include "lenia_hip_module.h"

#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>

using life_hip::DeviceImage;
using life_hip::SimulationContext;

LeniaHipModule::~LeniaHipModule() {
    if (rgba_) (void)hipFree(rgba_);
    if (world_) lenia_world_destroy(world_);
    if (ownsCtx_ && ctx_) lenia_ctx_destroy(ctx_);
}

void LeniaHipModule::report(const char *what) {
    err_ = std::string(what) + ": " + lenia_last_error();
    std::fprintf(stderr, "[life] %s: %s\n", instanceName_.c_str(), err_.c_str());
}

int LeniaHipModule::resolve(const std::string &instanceName, const std::string &paramsJson,
                            uint32_t W, uint32_t H, lenia_world_desc *desc, lenia_colormap *cmap) {
    // a one-module scene around the params: lenia_scene_describe applies
    // LeniaModule::setup's resolution to it
    const std::string scene = "{\"resolution\": [" + std::to_string(W) + ", " + std::to_string(H) +
                              "], \"modules\": [{\"type\": \"Lenia\", \"name\": \"" + instanceName +
                              "\", \"enabled\": true, \"params\": " + paramsJson + "}]}";
    char name[128];
    return lenia_scene_describe(scene.data(), scene.size(), 0, desc, cmap, name, sizeof(name));
}

void LeniaHipModule::setOrganism(std::vector<float> cells, uint32_t width, uint32_t height,
                                 uint32_t stampCount, bool stampRotate) {
    organism_ = std::move(cells);
    orgW_ = width;
    orgH_ = height;
    stampCount_ = stampCount;
    stampRotate_ = stampRotate;
    if (world_ && !organism_.empty() &&
        lenia_world_set_organism(world_, organism_.data(), orgW_, orgH_, stampCount_, stampRotate_ ? 1u : 0u))
        report("organism");
}

// LeniaModule::setup (Lenia.cpp:208-408)
void LeniaHipModule::setup(SimulationContext &ctx) {
    err_.clear();
    if (resolve(instanceName_, params_.text, ctx.width, ctx.height, &setup_, &cmap_)) {
        report("setup");
        return;
    }
    implicitKernel_ = setup_.convMode == LENIA_CONV_FFT && !params_.contains("kernels");
    if (!ctx_ && lenia_ctx_create(0, &ctx_)) {
        report("context");
        return;
    }
    if (lenia_world_create(ctx_, &setup_, &world_)) {
        report("world");
        world_ = nullptr;
        return;
    }
    if (!organism_.empty() &&
        lenia_world_set_organism(world_, organism_.data(), orgW_, orgH_, stampCount_, stampRotate_ ? 1u : 0u)) {
        report("organism");
        return;
    }
    outW_ = ctx.width;
    outH_ = ctx.height;
    if (hipMalloc((void **)&rgba_, sizeof(uint16_t) * 4 * outW_ * outH_) != hipSuccess) {
        rgba_ = nullptr;
        err_ = "output texture: device allocation failed";
        std::fprintf(stderr, "[life] %s: %s\n", instanceName_.c_str(), err_.c_str());
        return;
    }
    cur_ = setup_;
    needsInit_ = true;
}

// LeniaModule::reset (Lenia.cpp:448-451): seed_ = deriveSeed(seed, 'LEN1'),
// init at the next encode
void LeniaHipModule::reset(uint32_t seed) {
    if (!ok()) return;
    if (lenia_world_reset(world_, seed)) report("reset");
    needsInit_ = true;
}

// the per-frame reads of LeniaModule::encode (Lenia.cpp:456-464); the values
// resolved at setup are the fallbacks (the scene's own value when it has one)
void LeniaHipModule::frameParams(const SimulationContext &ctx, lenia_world_desc &d) const {
    d.dt = param(ctx, "dt", setup_.dt);
    d.noiseAmount = param(ctx, "noiseAmount", setup_.noiseAmount);
    d.muJitter = param(ctx, "muJitter", setup_.muJitter);
    d.injectAmount = param(ctx, "injectAmount", setup_.injectAmount);
    d.injectCount = (uint32_t)(param(ctx, "injectCount", (float)setup_.injectCount) + 0.5f);
    // the shaders test growthMode == 1; every other value is mode 0
    d.growthMode = (uint32_t)(param(ctx, "growthMode", (float)setup_.growthMode) + 0.5f) == 1u ? 1u : 0u;
    if (d.convMode != LENIA_CONV_FFT || (needsInit_ && implicitKernel_)) {
        // direct: leniaStep reads growthMu / growthSigma every frame; FFT with
        // the implicit kernel: they become its mu / sigma at init only
        d.kernels[0].mu = param(ctx, "growthMu", setup_.kernels[0].mu);
        d.kernels[0].sigma = param(ctx, "growthSigma", setup_.kernels[0].sigma);
    }
}

// LeniaModule::encode (Lenia.cpp:455-653)
void LeniaHipModule::encode(SimulationContext &ctx) {
    if (!ok()) return;
    lenia_world_desc d = cur_;
    frameParams(ctx, d);
    if (std::memcmp(&d, &cur_, sizeof(d)) != 0) {
        if (lenia_world_set_params(world_, &d)) {
            report("parameters");
            return;
        }
        cur_ = d;
    }
    if (lenia_world_frame(world_, ctx.frameIndex, ctx.substeps)) {
        report("frame");
        return;
    }
    needsInit_ = false;
    float *state = nullptr;
    if (lenia_world_state_device(world_, &state) ||
        lenia_colormap_field(ctx_, state, cur_.width, cur_.height, &cmap_, rgba_, outW_, outH_))
        report("colour map");
}

DeviceImage LeniaHipModule::outputTexture() const {
    DeviceImage im;
    if (rgba_) {
        im.data = rgba_;
        im.width = outW_;
        im.height = outH_;
        im.channels = 4;
    }
    return im;
}

DeviceImage LeniaHipModule::outputField() const {
    DeviceImage im;
    float *state = nullptr;
    if (world_ && lenia_world_state_device(world_, &state) == LENIA_OK) {
        im.data = state;
        im.width = cur_.width;
        im.height = cur_.height;
        im.channels = 1;
    }
    return im;
}
