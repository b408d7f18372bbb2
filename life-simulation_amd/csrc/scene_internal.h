// Scene JSON -> Lenia worlds: the parts of the reference's scene loader and
// LeniaModule::setup that decide what a world computes.  Shared by the scene
// runner and the sweep driver.
#pragma once

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "host_internal.h"
#include "json_lite.h"

namespace lenia {

// one "modules" entry (LifeCore/Params/Scene.cpp:36-47)
struct SceneModuleSpec {
    std::string type, name;
    bool enabled = true;
    json::Value params;
};

// Scene::parse (Scene.cpp:24-113): seed, resolution and module list
struct SceneSpec {
    uint32_t seed = 1234, width = 1920, height = 1080;
    std::vector<SceneModuleSpec> modules;
};
bool scene_parse(const json::Value &j, SceneSpec &out, std::string &err);

// organism file (loadOrganismJSON, Modules/FieldModules/Lenia.cpp:137-204)
struct Organism {
    float R = 13.0f, T = 10.0f, mu = 0.15f, sigma = 0.015f;
    uint32_t width = 0, height = 0;
    std::vector<float> cells;
};
bool organism_load(const std::string &path, Organism &out, std::string &err);

// What LeniaModule::setup resolves from a module's params (Lenia.cpp:208-356):
// the world descriptor, the palette, the per-frame fallbacks, and whether the
// FFT form's single implicit kernel follows growthMu / growthSigma at init.
struct LeniaSetup {
    std::string name;
    lenia_world_desc desc{};
    ColorMapParams cmap{};
    float opacity = 1.0f;
    bool fft = false;
    bool implicitKernel = false;  // FFT without "kernels"
    float fallbackMu = 0.15f, fallbackSigma = 0.017f, fallbackDt = 0.1f;
    bool stamp = false;  // stamp init with a loaded organism
    Organism organism;
    uint32_t stampCount = 6;
    bool stampRotate = true;
    json::Value params;
};
// Warnings the reference prints during setup go to stderr the same way.
bool lenia_setup(const SceneModuleSpec &m, uint32_t sceneW, uint32_t sceneH, LeniaSetup &out,
                 std::string &err);

// The values a module's encode() reads from the ParameterBus each frame
// (Lenia.cpp:456-464): scene-JSON numbers (booleans as 0 / 1) of the module's
// params, overridden by values set at run time, else the fallback.
float module_param(const LeniaSetup &m, const std::map<std::string, float> &bus,
                   const std::string &key, float fallback);
// the per-frame fields of the descriptor from those values; at init also
// the implicit FFT kernel's growth centre and width
void apply_frame_params(const LeniaSetup &m, const std::map<std::string, float> &bus, bool atInit,
                        lenia_world_desc &d);

void colormap_lenia_defaults(ColorMapParams &p);

}  // namespace lenia
