// Scene runner for Lenia scenes and the colour-map entry point.
//
// scene_parse follows Scene::parse (LifeCore/Params/Scene.cpp:24-113) for the
// fields a Lenia scene uses; lenia_setup follows LeniaModule::setup
// (Modules/FieldModules/Lenia.cpp:208-356) including its fallbacks and the
// warnings it prints; the runner steps modules frame by frame with the
// visibility / reseed / freeze rules of SceneRunner::step
// (LifeCore/Sim/SceneRunner.cpp:146-215) and colour-maps each encoded module
// into its RGBA16F layer (Lenia.cpp:646-652).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

#include "host_internal.h"
#include "launch.h"
#include "scene_internal.h"

using namespace lenia;

namespace {

bool is_pow2(uint32_t v) { return v && !(v & (v - 1)); }

// nlohmann's get<uint32_t>() of a JSON number: a plain conversion
uint32_t to_u32(double v) {
    if (!(v >= 0.0)) return 0u;
    if (v >= 4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)v;
}

void warn(const std::string &name, const std::string &msg) {
    std::fprintf(stderr, "[lenia] %s: %s\n", name.c_str(), msg.c_str());
}

}  // namespace

void lenia::colormap_lenia_defaults(ColorMapParams &p) {
    // ColorMapParams defaults (ColorMapPass.h:16-25), then Lenia's own palette
    // phase, channel and input scale (Lenia.cpp:350-354)
    for (int c = 0; c < 3; ++c) {
        p.a[c] = 0.5f;
        p.b[c] = 0.5f;
        p.c[c] = 1.0f;
    }
    p.d[0] = 0.60f;
    p.d[1] = 0.45f;
    p.d[2] = 0.30f;
    p.inputScale = 1.0f;
    p.inputBias = 0.0f;
    p.exposure = 1.0f;
    p.channel = 0u;
}

bool lenia::scene_parse(const json::Value &j, SceneSpec &out, std::string &err) {
    out = SceneSpec();
    if (!j.is_object()) {
        err = "scene JSON must be an object";
        return false;
    }
    double seed;
    if (!json::get_number(j, "seed", 1234.0, seed, err)) return false;
    out.seed = to_u32(seed);
    const json::Value *res = j.get("resolution");
    if (res && res->is_array() && res->arr.size() == 2) {
        if (!res->arr[0].is_number() || !res->arr[1].is_number()) {
            err = "\"resolution\" entries must be numbers";
            return false;
        }
        out.width = to_u32(res->arr[0].num);
        out.height = to_u32(res->arr[1].num);
    }
    const json::Value *mods = j.get("modules");
    if (mods && !mods->is_array()) {
        err = "\"modules\" must be an array";
        return false;
    }
    if (mods) {
        for (const json::Value &m : mods->arr) {
            SceneModuleSpec spec;
            if (!m.is_object()) {
                err = "module entry must be an object";
                return false;
            }
            if (!json::get_string(m, "type", "", spec.type, err)) return false;
            if (!json::get_string(m, "name", spec.type, spec.name, err)) return false;
            if (!json::get_bool(m, "enabled", true, spec.enabled, err)) return false;
            const json::Value *p = m.get("params");
            if (p) {
                if (!p->is_object()) {
                    err = "module \"params\" must be an object";
                    return false;
                }
                spec.params = *p;
            } else {
                spec.params.kind = json::Value::Object;
            }
            if (spec.type.empty()) {
                err = "module entry missing \"type\"";
                return false;
            }
            out.modules.push_back(std::move(spec));
        }
    }
    return true;
}

bool lenia::organism_load(const std::string &path, Organism &out, std::string &err) {
    std::ifstream f(path, std::ios::binary);
    if (!f) {
        err = "cannot open organism file: " + path;
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string text = ss.str();
    json::Value j;
    std::string perr;
    if (!json::parse(text.data(), text.size(), j, perr)) {
        err = "organism JSON parse error: " + perr;
        return false;
    }
    const json::Value *cells = j.get("cells");
    if (!cells || !cells->is_array() || cells->arr.empty()) {
        err = "organism JSON missing non-empty \"cells\" array";
        return false;
    }
    const size_t rows = cells->arr.size();
    if (!cells->arr[0].is_array() || cells->arr[0].arr.empty()) {
        err = "organism JSON \"cells\" rows must be non-empty arrays";
        return false;
    }
    const size_t cols = cells->arr[0].arr.size();
    out = Organism();
    out.cells.reserve(rows * cols);
    for (size_t r = 0; r < rows; ++r) {
        const json::Value &row = cells->arr[r];
        if (!row.is_array() || row.arr.size() != cols) {
            err = "organism JSON \"cells\" row " + std::to_string(r) + " length mismatch (expected " +
                  std::to_string(cols) + " columns)";
            return false;
        }
        for (size_t c = 0; c < cols; ++c) {
            if (!row.arr[c].is_number()) {
                err = "organism JSON cell at row " + std::to_string(r) + " col " + std::to_string(c) +
                      " is not a number";
                return false;
            }
            const float v = (float)row.arr[c].num;
            if (v < 0.0f || v > 1.0f) {
                err = "organism JSON cell value out of [0,1] at row " + std::to_string(r) + " col " +
                      std::to_string(c);
                return false;
            }
            out.cells.push_back(v);
        }
    }
    double R, T, mu, sigma;
    if (!json::get_number(j, "R", 13.0, R, err) || !json::get_number(j, "T", 10.0, T, err) ||
        !json::get_number(j, "mu", 0.15, mu, err) || !json::get_number(j, "sigma", 0.015, sigma, err))
        return false;
    out.R = (float)R;
    out.T = (float)T;
    out.mu = (float)mu;
    out.sigma = (float)sigma;
    out.width = (uint32_t)cols;
    out.height = (uint32_t)rows;
    return true;
}

float lenia::module_param(const LeniaSetup &m, const std::map<std::string, float> &bus,
                          const std::string &key, float fallback) {
    auto it = bus.find(m.name + "." + key);
    if (it != bus.end()) return it->second;
    const json::Value *v = m.params.get(key);
    if (v && v->is_number()) return (float)v->num;
    if (v && v->is_bool()) return v->b ? 1.0f : 0.0f;
    return fallback;
}

void lenia::apply_frame_params(const LeniaSetup &m, const std::map<std::string, float> &bus, bool atInit,
                               lenia_world_desc &d) {
    d.dt = module_param(m, bus, "dt", m.fallbackDt);
    const float growthMu = module_param(m, bus, "growthMu", m.fallbackMu);
    const float growthSigma = module_param(m, bus, "growthSigma", m.fallbackSigma);
    d.noiseAmount = module_param(m, bus, "noiseAmount", 0.0f);
    d.muJitter = module_param(m, bus, "muJitter", 0.0f);
    d.injectAmount = module_param(m, bus, "injectAmount", 0.0f);
    d.injectCount = to_u32(module_param(m, bus, "injectCount", 3.0f) + 0.5f);
    // the shaders test growthMode == 1; every other value is mode 0
    d.growthMode = to_u32(module_param(m, bus, "growthMode", 1.0f) + 0.5f) == 1u ? 1u : 0u;
    if (!m.fft || (atInit && m.implicitKernel)) {
        // direct: leniaStep reads growthMu / growthSigma every frame; FFT with
        // the implicit kernel: they become kMu / kSigma at init only
        d.kernels[0].mu = growthMu;
        d.kernels[0].sigma = growthSigma;
    }
}

bool lenia::lenia_setup(const SceneModuleSpec &m, uint32_t sceneW, uint32_t sceneH, LeniaSetup &out,
                        std::string &err) {
    out = LeniaSetup();
    out.name = m.name;
    out.params = m.params;
    const json::Value &P = m.params;
    std::string convMode;
    if (!json::get_string(P, "convMode", "direct", convMode, err)) return false;
    bool useFFT = convMode == "fft";
    double v;
    if (!json::get_number(P, "simWidth", 0.0, v, err)) return false;
    const uint32_t simW = to_u32(v) ? to_u32(v) : sceneW;
    if (!json::get_number(P, "simHeight", 0.0, v, err)) return false;
    const uint32_t simH = to_u32(v) ? to_u32(v) : sceneH;

    struct KernelDef {
        float radiusScale = 1.0f, mu = 0.15f, sigma = 0.017f, weight = 1.0f;
        std::vector<float> betas;
    };
    std::vector<KernelDef> kernels;
    const json::Value *kj = P.get("kernels");
    const bool hasKernels = kj && kj->is_array() && !kj->arr.empty();
    if (hasKernels) {
        for (const json::Value &k : kj->arr) {
            if (kernels.size() >= 4) {
                warn(out.name, "\"kernels\" has more than 4 entries; truncating to 4");
                break;
            }
            if (!k.is_object()) {
                err = "\"kernels\" entries must be objects";
                return false;
            }
            KernelDef kd;
            double a, b, c, d;
            if (!json::get_number(k, "radiusScale", 1.0, a, err) || !json::get_number(k, "mu", 0.15, b, err) ||
                !json::get_number(k, "sigma", 0.017, c, err) || !json::get_number(k, "weight", 1.0, d, err))
                return false;
            kd.radiusScale = (float)a;
            kd.mu = (float)b;
            kd.sigma = (float)c;
            kd.weight = (float)d;
            const json::Value *betas = k.get("betas");
            if (betas && betas->is_array()) {
                for (const json::Value &bv : betas->arr) {
                    if (!bv.is_number()) {
                        err = "\"betas\" entries must be numbers";
                        return false;
                    }
                    kd.betas.push_back((float)bv.num);
                }
            }
            if (kd.betas.empty()) kd.betas.push_back(1.0f);
            if (kd.betas.size() > 4) {
                err = "kernels with more than 4 betas are not supported";
                return false;
            }
            kernels.push_back(kd);
        }
    }
    if (useFFT && (!is_pow2(simW) || !is_pow2(simH))) {
        warn(out.name, "convMode=\"fft\" requires simWidth/simHeight to be powers of two (got " +
                           std::to_string(simW) + "x" + std::to_string(simH) +
                           "); falling back to convMode=\"direct\"");
        useFFT = false;
    }
    float radiusScaleOverride = 1.0f;
    if (hasKernels && !useFFT) {
        warn(out.name, "\"kernels\" requires convMode=\"fft\"; continuing in direct mode using kernels[0] only");
        radiusScaleOverride = kernels[0].radiusScale;
        out.fallbackMu = kernels[0].mu;
        out.fallbackSigma = kernels[0].sigma;
    }
    if (!useFFT) kernels.clear();

    // organism + stamp init (Lenia.cpp:274-326)
    std::string initMode, organismPath;
    bool wantOrganismParams, stampRotate;
    if (!json::get_string(P, "initMode", "noise", initMode, err) ||
        !json::get_string(P, "organism", "", organismPath, err) ||
        !json::get_bool(P, "useOrganismParams", true, wantOrganismParams, err) ||
        !json::get_bool(P, "stampRotate", true, stampRotate, err) ||
        !json::get_number(P, "stampCount", 6.0, v, err))
        return false;
    bool wantStamp = initMode == "stamp";
    out.stampCount = std::min(32u, to_u32(v));
    out.stampRotate = stampRotate;
    if (!organismPath.empty() && wantOrganismParams && hasKernels) {
        warn(out.name, "\"useOrganismParams\" cannot be combined with \"kernels\"; ignoring organism parameter overrides");
        wantOrganismParams = false;
    }
    if (wantStamp && organismPath.empty()) {
        warn(out.name, "initMode=\"stamp\" requires an \"organism\" path; falling back to noise init");
        wantStamp = false;
    }
    uint32_t radiusDefault = 13u;
    bool organismLoaded = false;
    if (!organismPath.empty()) {
        std::string lerr;
        if (organism_load(organismPath, out.organism, lerr)) {
            organismLoaded = true;
        } else {
            warn(out.name, "failed to load organism \"" + organismPath + "\": " + lerr +
                               "; falling back to noise init");
            wantStamp = false;
        }
    }
    if (organismLoaded && wantOrganismParams) {
        radiusDefault = (uint32_t)std::lround(out.organism.R);
        out.fallbackMu = out.organism.mu;
        out.fallbackSigma = out.organism.sigma;
        if (out.organism.T > 1e-6f) out.fallbackDt = 1.0f / out.organism.T;
    }
    out.stamp = wantStamp && organismLoaded;

    // radius and kernel-shell parameters (Lenia.cpp:339-347)
    if (!json::get_number(P, "radius", (double)radiusDefault, v, err)) return false;
    const uint32_t R = (uint32_t)std::lround((float)to_u32(v) * radiusScaleOverride);
    double cov, scale, shellMu, shellSigma, opacity;
    if (!json::get_number(P, "initCoverage", 0.4, cov, err) || !json::get_number(P, "initScale", 24.0, scale, err) ||
        !json::get_number(P, "kernelShellMu", 0.5, shellMu, err) ||
        !json::get_number(P, "kernelShellSigma", 0.15, shellSigma, err) ||
        !json::get_number(P, "opacity", 1.0, opacity, err))
        return false;
    out.opacity = (float)opacity;
    std::string core;
    if (!json::get_string(P, "kernelCore", "gaussian", core, err)) return false;
    if (core != "gaussian" && core != "exp") {
        err = "\"kernelCore\" must be \"gaussian\" or \"exp\"";
        return false;
    }

    // palette (ColorMapPass::configure, ColorMapPass.cpp:28-100); relief
    // shading is another renderer, not part of this library
    colormap_lenia_defaults(out.cmap);
    if (P.contains("relief")) warn(out.name, "\"relief\" shading is not supported; using the palette");
    const json::Value *cm = P.get("colorMap");
    if (cm && cm->is_object()) {
        std::string palette;
        if (!json::get_string(*cm, "palette", "", palette, err)) return false;
        if (palette == "mono")
            for (int c = 0; c < 3; ++c) out.cmap.a[c] = out.cmap.b[c] = out.cmap.c[c] = out.cmap.d[c] = 0.5f;
        auto vec3 = [&](const char *key, float *dst) -> bool {
            const json::Value *a = cm->get(key);
            if (a && a->is_array() && a->arr.size() >= 3) {
                for (int c = 0; c < 3; ++c) {
                    if (!a->arr[c].is_number()) {
                        err = std::string("colorMap.") + key + " entries must be numbers";
                        return false;
                    }
                    dst[c] = (float)a->arr[c].num;
                }
            }
            return true;
        };
        if (!vec3("a", out.cmap.a) || !vec3("b", out.cmap.b) || !vec3("c", out.cmap.c) || !vec3("d", out.cmap.d))
            return false;
        double s, b, e, ch;
        if (!json::get_number(*cm, "inputScale", out.cmap.inputScale, s, err) ||
            !json::get_number(*cm, "inputBias", out.cmap.inputBias, b, err) ||
            !json::get_number(*cm, "exposure", out.cmap.exposure, e, err) ||
            !json::get_number(*cm, "channel", out.cmap.channel, ch, err))
            return false;
        out.cmap.inputScale = (float)s;
        out.cmap.inputBias = (float)b;
        out.cmap.exposure = (float)e;
        out.cmap.channel = to_u32(ch);
    }

    // the world descriptor
    lenia_world_desc &d = out.desc;
    std::memset(&d, 0, sizeof(d));
    d.width = simW;
    d.height = simH;
    d.convMode = useFFT ? LENIA_CONV_FFT : LENIA_CONV_DIRECT;
    d.radius = R;
    out.fft = useFFT;
    const uint32_t coreId = core == "exp" ? LENIA_CORE_EXP : LENIA_CORE_GAUSSIAN;
    if (useFFT) {
        out.implicitKernel = kernels.empty();
        if (kernels.empty()) kernels.push_back(KernelDef{1.0f, 0.15f, 0.017f, 1.0f, {1.0f}});
        d.nKernels = (uint32_t)kernels.size();
        for (size_t k = 0; k < kernels.size(); ++k) {
            lenia_kernel_desc &kd = d.kernels[k];
            kd.shape = LENIA_SHAPE_RING;
            kd.core = coreId;
            kd.radius = (float)R * kernels[k].radiusScale;
            kd.shellMu = (float)shellMu;
            kd.shellSigma = (float)shellSigma;
            kd.nBetas = (uint32_t)kernels[k].betas.size();
            for (size_t b = 0; b < kernels[k].betas.size(); ++b) kd.betas[b] = kernels[k].betas[b];
            kd.mu = kernels[k].mu;
            kd.sigma = kernels[k].sigma;
            kd.weight = kernels[k].weight;
        }
    } else {
        d.nKernels = 1;
        lenia_kernel_desc &kd = d.kernels[0];
        kd.shape = LENIA_SHAPE_SHELL;
        kd.core = coreId;
        kd.radius = (float)R;
        kd.shellMu = (float)shellMu;
        kd.shellSigma = (float)shellSigma;
        kd.nBetas = 1;
        kd.betas[0] = 1.0f;
        kd.weight = 1.0f;
    }
    d.initMode = out.stamp ? LENIA_INIT_STAMP : LENIA_INIT_NOISE;
    d.initCoverage = (float)cov;
    d.initScale = (float)scale;
    apply_frame_params(out, {}, true, d);
    return true;
}

// ------------------------------------------------------------------ colour map

extern "C" void lenia_colormap_lenia_defaults(lenia_colormap *p) {
    if (!p) return;
    ColorMapParams c;
    colormap_lenia_defaults(c);
    static_assert(sizeof(ColorMapParams) == sizeof(lenia_colormap), "palette layouts differ");
    std::memcpy(p, &c, sizeof(c));
}

extern "C" int lenia_colormap_field(lenia_ctx *ctx, const float *field, uint32_t srcW, uint32_t srcH,
                                    const lenia_colormap *p, uint16_t *rgba, uint32_t dstW, uint32_t dstH) {
    if (!ctx || !field || !p || !rgba) return fail(LENIA_ERR_INVALID_ARG, "lenia_colormap_field: NULL argument");
    if (srcW < 1 || srcH < 1 || dstW < 1 || dstH < 1 || srcW > 65536 || srcH > 65536 || dstW > 65536 ||
        dstH > 65536)
        return fail(LENIA_ERR_INVALID_ARG, "lenia_colormap_field: size out of range");
    if (((uintptr_t)rgba & 7u) != 0) return fail(LENIA_ERR_INVALID_ARG, "lenia_colormap_field: rgba not 8-byte aligned");
    (void)hipSetDevice(ctx->device);
    ColorMapParams c;
    std::memcpy(&c, p, sizeof(c));
    hipError_t e = launch_colormap(field, (int)srcW, (int)srcH, c, rgba, (int)dstW, (int)dstH, ctx->stream);
    return e == hipSuccess ? LENIA_OK : hip_fail(e, "colour map");
}

// ------------------------------------------------------------------ scene runner

struct lenia_scene {
    lenia_ctx *ctx = nullptr;
    SceneSpec spec;
    uint32_t substeps = 1;
    uint32_t frameIndex = 0;
    struct Mod {
        LeniaSetup setup;
        lenia_world *world = nullptr;
        lenia_world_desc cur{};
        uint16_t *layer = nullptr;
        bool visibleWasHigh = false, reseedWasHigh = false, encodedOnce = false, needsInit = true;
    };
    std::vector<Mod> mods;
    std::map<std::string, float> bus;
};

namespace {

int parse_text(const char *json, size_t len, json::Value &j) {
    std::string err;
    if (!json || !json::parse(json, len, j, err)) return fail(LENIA_ERR_INVALID_ARG, "scene JSON parse error: %s", err.c_str());
    return LENIA_OK;
}

// the enabled modules of a scene, all of type "Lenia"
int lenia_modules(const SceneSpec &spec, std::vector<const SceneModuleSpec *> &out) {
    out.clear();
    for (const SceneModuleSpec &m : spec.modules) {
        if (!m.enabled) continue;
        if (m.type != "Lenia")
            return fail(LENIA_ERR_UNSUPPORTED, "module \"%s\" has type \"%s\"; only \"Lenia\" modules are supported",
                        m.name.c_str(), m.type.c_str());
        out.push_back(&m);
    }
    if (out.empty()) return fail(LENIA_ERR_INVALID_ARG, "scene contains no enabled modules");
    return LENIA_OK;
}

void scene_release(lenia_scene *s) {
    for (auto &m : s->mods) {
        if (m.world) lenia_world_destroy(m.world);
        dev_free(s->ctx, m.layer);
        m.world = nullptr;
        m.layer = nullptr;
    }
}

}  // namespace

extern "C" int lenia_scene_describe(const char *json, size_t len, uint32_t module, lenia_world_desc *desc,
                                    lenia_colormap *cmap, char *name, size_t nameCap) {
    if (!json || !desc) return fail(LENIA_ERR_INVALID_ARG, "lenia_scene_describe: NULL argument");
    json::Value j;
    int rc = parse_text(json, len, j);
    if (rc) return rc;
    SceneSpec spec;
    std::string err;
    if (!scene_parse(j, spec, err)) return fail(LENIA_ERR_INVALID_ARG, "scene JSON structure error: %s", err.c_str());
    std::vector<const SceneModuleSpec *> mods;
    rc = lenia_modules(spec, mods);
    if (rc) return rc;
    if (module >= mods.size()) return fail(LENIA_ERR_INVALID_ARG, "module index %u out of range", module);
    LeniaSetup setup;
    if (!lenia_setup(*mods[module], spec.width, spec.height, setup, err))
        return fail(LENIA_ERR_INVALID_ARG, "module \"%s\": %s", mods[module]->name.c_str(), err.c_str());
    if (desc) *desc = setup.desc;
    if (cmap) std::memcpy(cmap, &setup.cmap, sizeof(*cmap));
    if (name && nameCap) {
        std::snprintf(name, nameCap, "%s", setup.name.c_str());
    }
    return LENIA_OK;
}

extern "C" int lenia_scene_create(lenia_ctx *ctx, const char *json, size_t len, uint32_t substeps,
                                  lenia_scene **out) {
    if (!ctx || !out) return fail(LENIA_ERR_INVALID_ARG, "lenia_scene_create: NULL argument");
    *out = nullptr;
    json::Value j;
    int rc = parse_text(json, len, j);
    if (rc) return rc;
    lenia_scene *s = new lenia_scene();
    s->ctx = ctx;
    s->substeps = substeps;
    std::string err;
    if (!scene_parse(j, s->spec, err)) {
        delete s;
        return fail(LENIA_ERR_INVALID_ARG, "scene JSON structure error: %s", err.c_str());
    }
    if (s->spec.width < 1 || s->spec.height < 1 || s->spec.width > 16384 || s->spec.height > 16384) {
        delete s;
        return fail(LENIA_ERR_INVALID_ARG, "scene resolution out of range");
    }
    std::vector<const SceneModuleSpec *> mods;
    rc = lenia_modules(s->spec, mods);
    if (rc) {
        delete s;
        return rc;
    }
    (void)hipSetDevice(ctx->device);
    s->mods.resize(mods.size());
    for (size_t i = 0; i < mods.size() && !rc; ++i) {
        lenia_scene::Mod &m = s->mods[i];
        if (!lenia_setup(*mods[i], s->spec.width, s->spec.height, m.setup, err)) {
            rc = fail(LENIA_ERR_INVALID_ARG, "module \"%s\": %s", mods[i]->name.c_str(), err.c_str());
            break;
        }
        m.cur = m.setup.desc;
        rc = lenia_world_create(ctx, &m.cur, &m.world);
        if (!rc && m.setup.stamp)
            rc = lenia_world_set_organism(m.world, m.setup.organism.cells.data(), m.setup.organism.width,
                                          m.setup.organism.height, m.setup.stampCount,
                                          m.setup.stampRotate ? 1u : 0u);
        if (!rc)
            rc = dev_alloc(ctx, (void **)&m.layer, (size_t)s->spec.width * s->spec.height * 4 * sizeof(uint16_t));
        if (!rc) {
            // a layer no frame has encoded yet is transparent black
            hipError_t e = launch_fill(reinterpret_cast<float *>(m.layer), (size_t)s->spec.width * s->spec.height * 2,
                                       0.0f, ctx->stream);
            if (e != hipSuccess) rc = hip_fail(e, "clear layer");
        }
    }
    if (!rc) rc = lenia_scene_reset(s);
    if (rc) {
        scene_release(s);
        delete s;
        return rc;
    }
    *out = s;
    return LENIA_OK;
}

extern "C" int lenia_scene_destroy(lenia_scene *s) {
    if (!s) return LENIA_OK;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    scene_release(s);
    delete s;
    return LENIA_OK;
}

extern "C" int lenia_scene_reset(lenia_scene *s) {
    if (!s) return fail(LENIA_ERR_INVALID_ARG, "lenia_scene_reset: NULL scene");
    s->frameIndex = 0;
    for (auto &m : s->mods) {
        int rc = lenia_world_reset(m.world, s->spec.seed);
        if (rc) return rc;
        m.needsInit = true;
        m.visibleWasHigh = m.reseedWasHigh = m.encodedOnce = false;
    }
    return LENIA_OK;
}

// one SceneRunner::step for the Lenia modules (SceneRunner.cpp:146-215)
static int scene_frame(lenia_scene *s) {
    lenia_ctx *ctx = s->ctx;
    for (auto &m : s->mods) {
        const LeniaSetup &st = m.setup;
        const float op = module_param(st, s->bus, "opacity", st.opacity);
        const bool visible = op > 0.02f;  // SceneRunner::VISIBLE_EPS
        const bool appearing = visible && !m.visibleWasHigh;
        m.visibleWasHigh = visible;
        if (!visible) continue;
        const bool reseedHigh = module_param(st, s->bus, "reseed", 0.0f) > 0.5f;
        if (appearing || (reseedHigh && !m.reseedWasHigh)) {
            int rc = lenia_world_reset(m.world, s->spec.seed ^ (s->frameIndex * 2654435761u));
            if (rc) return rc;
            m.needsInit = true;
            m.encodedOnce = false;
        }
        m.reseedWasHigh = reseedHigh;
        if (m.encodedOnce && module_param(st, s->bus, "freeze", 0.0f) > 0.5f) continue;
        // LeniaModule::encode: this frame's parameters, init if pending, the
        // substeps, then the colour layer
        lenia_world_desc d = m.cur;
        apply_frame_params(st, s->bus, m.needsInit, d);
        if (std::memcmp(&d, &m.cur, sizeof(d)) != 0) {
            int rc = lenia_world_set_params(m.world, &d);
            if (rc) return rc;
            m.cur = d;
        }
        int rc = lenia_world_frame(m.world, s->frameIndex, s->substeps);
        if (rc) return rc;
        m.needsInit = false;
        float *state = nullptr;
        rc = lenia_world_state_device(m.world, &state);
        if (rc) return rc;
        hipError_t e = launch_colormap(state, (int)d.width, (int)d.height, st.cmap, m.layer, (int)s->spec.width,
                                       (int)s->spec.height, ctx->stream);
        if (e != hipSuccess) return hip_fail(e, "colour map");
        m.encodedOnce = true;
    }
    s->frameIndex += 1;
    return LENIA_OK;
}

extern "C" int lenia_scene_step(lenia_scene *s, uint32_t nFrames) {
    if (!s) return fail(LENIA_ERR_INVALID_ARG, "lenia_scene_step: NULL scene");
    (void)hipSetDevice(s->ctx->device);
    lenia_ctx *ctx = s->ctx;
    // GPU time of the whole call (the frames' world steps and colour layers),
    // what LifeBench reports per frame (Apps/LifeBench/main.cpp:112-133)
    (void)hipEventRecord(ctx->sp0, ctx->stream);
    for (uint32_t f = 0; f < nFrames; ++f) {
        int rc = scene_frame(s);
        if (rc) return rc;
    }
    (void)hipEventRecord(ctx->sp1, ctx->stream);
    ctx->timed = true;
    ctx->last0 = ctx->sp0;
    ctx->last1 = ctx->sp1;
    return LENIA_OK;
}

extern "C" int lenia_scene_set_param(lenia_scene *s, const char *target, float value) {
    if (!s || !target) return fail(LENIA_ERR_INVALID_ARG, "lenia_scene_set_param: NULL argument");
    const char *dot = std::strchr(target, '.');
    if (!dot || dot == target || !dot[1]) return fail(LENIA_ERR_INVALID_ARG, "target \"%s\" must be \"module.param\"", target);
    s->bus[target] = value;
    return LENIA_OK;
}

extern "C" int lenia_scene_get_info(lenia_scene *s, lenia_scene_info *out) {
    if (!s || !out) return fail(LENIA_ERR_INVALID_ARG, "lenia_scene_get_info: NULL argument");
    out->width = s->spec.width;
    out->height = s->spec.height;
    out->seed = s->spec.seed;
    out->substeps = s->substeps;
    out->frameIndex = s->frameIndex;
    out->nModules = (uint32_t)s->mods.size();
    return LENIA_OK;
}

extern "C" int lenia_scene_world(lenia_scene *s, uint32_t module, lenia_world **world) {
    if (!s || !world) return fail(LENIA_ERR_INVALID_ARG, "lenia_scene_world: NULL argument");
    if (module >= s->mods.size()) return fail(LENIA_ERR_INVALID_ARG, "module index %u out of range", module);
    *world = s->mods[module].world;
    return LENIA_OK;
}

extern "C" int lenia_scene_layer(lenia_scene *s, uint32_t module, uint16_t *rgba, size_t count, int onDevice) {
    if (!s || !rgba) return fail(LENIA_ERR_INVALID_ARG, "lenia_scene_layer: NULL argument");
    if (module >= s->mods.size()) return fail(LENIA_ERR_INVALID_ARG, "module index %u out of range", module);
    const size_t n = (size_t)s->spec.width * s->spec.height * 4;
    if (count != n) return fail(LENIA_ERR_INVALID_ARG, "layer: count != width*height*4");
    (void)hipSetDevice(s->ctx->device);
    hipError_t e = hipMemcpyAsync(rgba, s->mods[module].layer, n * sizeof(uint16_t),
                                  onDevice ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s->ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->ctx->stream);
    return e == hipSuccess ? LENIA_OK : hip_fail(e, "read layer");
}
