// The part of the reference's module interface (LifeCore/Sim/SimulationModule.h:27-135)
// that a Lenia field module overrides, with the Metal types replaced by what
// the HIP library hands out: a texture handle becomes a device pointer with
// its size, the scene's nlohmann::json params become ModuleParams (the text of
// the module's "params" object and its keys).  Built inside the reference
// tree, LeniaHipModule derives from the real life::FieldModule instead and
// this header is not used.
#pragma once

#include <cstdint>
#include <set>
#include <string>

namespace life_hip {

// ParameterBus::value (LifeCore/Params/ParameterBus.h): the scene's base value
// of "module.key", plus whatever the audio mappings add this frame
struct ParameterBus {
    virtual ~ParameterBus() = default;
    virtual float value(const std::string &key, float fallback) const = 0;
};

// SimulationContext (SimulationModule.h:27-37) without the Metal members
struct SimulationContext {
    ParameterBus *params = nullptr;
    uint32_t frameIndex = 0;
    uint32_t substeps = 1;
    float dt = 1.0f / 60.0f;
    uint32_t width = 0;  // scene resolution
    uint32_t height = 0;
    uint32_t seed = 0;   // scene seed (Scene::seed)
};

struct ModuleParams {
    std::string text;           // the "params" object, as JSON text
    std::set<std::string> keys; // its top-level keys
    bool contains(const std::string &k) const { return keys.count(k) != 0; }
};

// a device image: `data` holds width x height texels of `channels` values
struct DeviceImage {
    const void *data = nullptr;
    uint32_t width = 0, height = 0, channels = 0;
    bool valid() const { return data != nullptr; }
};

// SimulationModule / FieldModule lifecycle (SimulationModule.h:39-135)
class FieldModule {
public:
    virtual ~FieldModule() = default;
    virtual void setup(SimulationContext &ctx) = 0;
    virtual void reset(uint32_t seed) = 0;
    virtual void updateCPU() {}
    virtual void encode(SimulationContext &ctx) = 0;
    virtual DeviceImage outputTexture() const = 0;  // RGBA16F colour layer
    virtual DeviceImage outputField() const = 0;    // R32F state field

    void configure(std::string instanceName, ModuleParams params) {
        instanceName_ = std::move(instanceName);
        params_ = std::move(params);
    }
    const std::string &instanceName() const { return instanceName_; }
    const ModuleParams &jsonParams() const { return params_; }

protected:
    float param(const SimulationContext &ctx, const std::string &key, float fallback) const {
        return ctx.params ? ctx.params->value(instanceName_ + "." + key, fallback) : fallback;
    }
    std::string instanceName_;
    ModuleParams params_;
};

}  // namespace life_hip
