// LeniaHipModule: the reference's Lenia field module (Modules/FieldModules/Lenia.h:50-160,
// Lenia.cpp:208-653) as a module class over the HIP library's C ABI
// (include/lenia_mi355x.h).  It is what a maintainer registers in place of
// LeniaModule (Modules/RegisterModules.cpp:23):
//
//   f.registerType("Lenia", [] { return std::make_unique<LeniaHipModule>(); });
//
// setup() resolves the module's params exactly as LeniaModule::setup does
// (lenia_scene_describe: simWidth/simHeight, convMode and its fallbacks,
// kernels, radius x radiusScale, palette) and creates the world; reset() and
// encode() follow LeniaModule::reset / encode: the per-frame values come from
// the ParameterBus with the set-up values as fallbacks, the first encode after
// a reset runs the init, then `substeps` steps with seed
// pcg(seed_ ^ (frame*197 + s)), then the colour map into the RGBA16F output.
// Errors are reported as LeniaModule reports its set-up problems (stderr,
// "[life] <instance>: ...") and leave the module inert.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "lenia_mi355x.h"
#include "module_api.h"

class LeniaHipModule : public life_hip::FieldModule {
public:
    // shared: a context owned by the caller (one per GPU, shared by modules);
    // NULL makes the module create and own one on device 0
    explicit LeniaHipModule(lenia_ctx *shared = nullptr) : ctx_(shared), ownsCtx_(shared == nullptr) {}
    ~LeniaHipModule() override;
    LeniaHipModule(const LeniaHipModule &) = delete;
    LeniaHipModule &operator=(const LeniaHipModule &) = delete;

    void setup(life_hip::SimulationContext &ctx) override;
    void reset(uint32_t seed) override;
    void encode(life_hip::SimulationContext &ctx) override;
    life_hip::DeviceImage outputTexture() const override;
    life_hip::DeviceImage outputField() const override;

    // organism cells for initMode "stamp" (row-major, values in [0, 1]); in the
    // reference tree setup() gets them from loadOrganismJSON (LifeCore)
    // (stampCount and stampRotate as LeniaModule::setup reads them)
    void setOrganism(std::vector<float> cells, uint32_t width, uint32_t height, uint32_t stampCount,
                     bool stampRotate);

    bool ok() const { return world_ != nullptr && err_.empty(); }
    const std::string &error() const { return err_; }
    lenia_ctx *context() const { return ctx_; }
    lenia_world *world() const { return world_; }
    const lenia_world_desc &descriptor() const { return cur_; }

    // the descriptor and palette setup() resolves for `params` at scene size
    // W x H (no GPU call)
    static int resolve(const std::string &instanceName, const std::string &paramsJson, uint32_t W,
                       uint32_t H, lenia_world_desc *desc, lenia_colormap *cmap);

private:
    void report(const char *what);
    void frameParams(const life_hip::SimulationContext &ctx, lenia_world_desc &d) const;

    lenia_ctx *ctx_ = nullptr;
    bool ownsCtx_ = true;
    lenia_world *world_ = nullptr;
    lenia_world_desc setup_{};  // values resolved at setup: the per-frame fallbacks
    lenia_world_desc cur_{};    // what the world runs with now
    lenia_colormap cmap_{};
    bool implicitKernel_ = false;  // FFT form without "kernels": growthMu/Sigma read at init only
    bool needsInit_ = true;
    uint16_t *rgba_ = nullptr;  // device, width x height RGBA16F
    uint32_t outW_ = 0, outH_ = 0;
    std::vector<float> organism_;
    uint32_t orgW_ = 0, orgH_ = 0, stampCount_ = 6;
    bool stampRotate_ = true;
    std::string err_;
};
