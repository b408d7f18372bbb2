// Parameter sweeps of a one-module Lenia scene on the batched launch.
//
// The reference's PresetLab (Apps/LifePresetLab/main.cpp) expands a sweep
// JSON into variants (resolveAxis :109-152, buildVariants :232-270, the
// 500-variant cap :678-687), applies each variant's overrides to a copy of
// the scene (applyOverrides :296-320) and runs one SceneRunner per variant,
// one after the other (:718-806), scoring the last frame's Rec.709 luma
// (:67-90).  Here every variant becomes one world and all of them step
// together as one batch per convolution form (worlds of the MFMA forms step
// alone); the colour layers and their luma statistics are computed on the GPU.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>

#include "host_internal.h"
#include "launch.h"
#include "scene_internal.h"

using namespace lenia;

struct lenia_sweep {
    lenia_ctx *ctx = nullptr;
    SceneSpec base;
    std::string mode;
    uint32_t substeps = 1;
    uint32_t framesRun = 0;
    struct Var {
        std::string id;
        uint32_t seed = 0;
        std::vector<std::pair<std::string, json::Value>> overrides;
        LeniaSetup setup;
        lenia_world *world = nullptr;
        uint16_t *layer = nullptr;
        double mean = 0.0, sd = 0.0;
        bool alive = false;
        bool visible = true, freeze = false;
        std::string error;
    };
    std::vector<Var> vars;
    std::vector<lenia_batch *> batches;
    std::vector<size_t> alone;  // variants stepped by themselves (MFMA forms)
    double *partial = nullptr, *stats = nullptr;
};

namespace {

constexpr size_t kMaxVariants = 500;

struct Axis {
    std::string target;
    std::vector<json::Value> values;
};

// resolveAxis (main.cpp:109-152)
bool resolve_axis(const json::Value &a, Axis &out, std::string &err) {
    if (!a.is_object()) {
        err = "sweep axes must be objects";
        return false;
    }
    if (!json::get_string(a, "target", "", out.target, err)) return false;
    if (out.target.empty()) {
        err = "sweep axis missing \"target\"";
        return false;
    }
    if (out.target.find('.') == std::string::npos) {
        err = "sweep axis target \"" + out.target + "\" must be \"module.param\"";
        return false;
    }
    if (const json::Value *vals = a.get("values")) {
        if (!vals->is_array() || vals->arr.empty()) {
            err = "axis \"" + out.target + "\": \"values\" must be a non-empty array";
            return false;
        }
        out.values = vals->arr;
        return true;
    }
    if (const json::Value *r = a.get("range")) {
        if (!r->is_array() || r->arr.size() != 2 || !r->arr[0].is_number() || !r->arr[1].is_number()) {
            err = "axis \"" + out.target + "\": \"range\" must be [lo, hi]";
            return false;
        }
        const double lo = r->arr[0].num, hi = r->arr[1].num;
        double sv;
        if (!json::get_number(a, "steps", 0.0, sv, err)) return false;
        const int steps = (int)sv;
        if (steps < 1) {
            err = "axis \"" + out.target + "\": \"steps\" must be >= 1";
            return false;
        }
        for (int i = 0; i < steps; ++i) {
            json::Value v;
            v.kind = json::Value::Number;
            // linear, both ends included
            v.num = steps == 1 ? lo : lo + (double)i / (double)(steps - 1) * (hi - lo);
            out.values.push_back(v);
        }
        return true;
    }
    err = "axis \"" + out.target + "\": need \"values\" or \"range\"+\"steps\"";
    return false;
}

void set_member(json::Value &obj, const std::string &key, const json::Value &v) {
    for (auto &m : obj.obj)
        if (m.first == key) {
            m.second = v;
            return;
        }
    obj.obj.emplace_back(key, v);
}

void sweep_release(lenia_sweep *s) {
    for (lenia_batch *b : s->batches) lenia_batch_destroy(b);
    s->batches.clear();
    for (auto &v : s->vars) {
        if (v.world) lenia_world_destroy(v.world);
        dev_free(s->ctx, v.layer);
        v.world = nullptr;
        v.layer = nullptr;
    }
    dev_free(s->ctx, s->partial);
    dev_free(s->ctx, s->stats);
    s->partial = s->stats = nullptr;
}

}  // namespace

extern "C" int lenia_sweep_create(lenia_ctx *ctx, const char *sceneJson, size_t sceneLen, const char *sweepJson,
                                  size_t sweepLen, uint32_t substeps, lenia_sweep **out) {
    if (!ctx || !sceneJson || !sweepJson || !out) return fail(LENIA_ERR_INVALID_ARG, "lenia_sweep_create: NULL argument");
    *out = nullptr;
    json::Value sj, wj;
    std::string err;
    if (!json::parse(sceneJson, sceneLen, sj, err)) return fail(LENIA_ERR_INVALID_ARG, "scene JSON parse error: %s", err.c_str());
    if (!json::parse(sweepJson, sweepLen, wj, err)) return fail(LENIA_ERR_INVALID_ARG, "sweep JSON parse error: %s", err.c_str());
    lenia_sweep *s = new lenia_sweep();
    s->ctx = ctx;
    s->substeps = substeps;
    auto bail = [&](int rc) {
        sweep_release(s);
        delete s;
        return rc;
    };
    if (!scene_parse(sj, s->base, err)) return bail(fail(LENIA_ERR_INVALID_ARG, "scene JSON structure error: %s", err.c_str()));
    const SceneModuleSpec *mod = nullptr;
    int enabled = 0;
    for (const SceneModuleSpec &m : s->base.modules) {
        if (!m.enabled) continue;
        ++enabled;
        mod = &m;
    }
    if (enabled != 1 || mod->type != "Lenia")
        return bail(fail(LENIA_ERR_UNSUPPORTED, "sweeps run scenes with exactly one enabled \"Lenia\" module"));
    if (s->base.width < 1 || s->base.height < 1 || s->base.width > 16384 || s->base.height > 16384)
        return bail(fail(LENIA_ERR_INVALID_ARG, "scene resolution out of range"));

    // loadSweep (main.cpp:154-195)
    if (!wj.is_object()) return bail(fail(LENIA_ERR_INVALID_ARG, "sweep JSON must be an object"));
    if (!json::get_string(wj, "mode", "cartesian", s->mode, err)) return bail(fail(LENIA_ERR_INVALID_ARG, "%s", err.c_str()));
    if (s->mode != "cartesian" && s->mode != "per-axis")
        return bail(fail(LENIA_ERR_INVALID_ARG, "sweep \"mode\" must be \"cartesian\" or \"per-axis\" (got \"%s\")",
                         s->mode.c_str()));
    std::vector<uint32_t> seeds;
    if (const json::Value *sv = wj.get("seeds")) {
        if (!sv->is_array()) return bail(fail(LENIA_ERR_INVALID_ARG, "sweep \"seeds\" must be an array"));
        for (const json::Value &x : sv->arr) {
            if (!x.is_number() || x.num < 0) return bail(fail(LENIA_ERR_INVALID_ARG, "sweep seeds must be unsigned numbers"));
            seeds.push_back((uint32_t)x.num);
        }
    }
    if (seeds.empty()) seeds.push_back(s->base.seed);
    std::vector<Axis> axes;
    if (const json::Value *av = wj.get("axes")) {
        if (!av->is_array()) return bail(fail(LENIA_ERR_INVALID_ARG, "sweep \"axes\" must be an array"));
        for (const json::Value &a : av->arr) {
            Axis ax;
            if (!resolve_axis(a, ax, err)) return bail(fail(LENIA_ERR_INVALID_ARG, "%s", err.c_str()));
            axes.push_back(std::move(ax));
        }
    }

    // buildVariants (main.cpp:232-270): cartesian odometer with the last axis
    // fastest, or one axis at a time; seeds outermost
    std::vector<std::pair<uint32_t, std::vector<std::pair<std::string, json::Value>>>> combos;
    if (s->mode == "cartesian") {
        std::vector<std::vector<size_t>> idx;
        bool empty = false;
        for (const Axis &a : axes) empty = empty || a.values.empty();
        if (!empty) {
            std::vector<size_t> c(axes.size(), 0);
            while (true) {
                idx.push_back(c);
                int a = (int)axes.size() - 1;
                while (a >= 0) {
                    if (++c[(size_t)a] < axes[(size_t)a].values.size()) break;
                    c[(size_t)a] = 0;
                    --a;
                }
                if (a < 0) break;
                if (idx.size() > kMaxVariants) break;
            }
        }
        for (uint32_t seed : seeds)
            for (const auto &ci : idx) {
                std::vector<std::pair<std::string, json::Value>> ov;
                for (size_t a = 0; a < axes.size(); ++a) ov.emplace_back(axes[a].target, axes[a].values[ci[a]]);
                combos.emplace_back(seed, std::move(ov));
                if (combos.size() > kMaxVariants) break;
            }
    } else {
        for (uint32_t seed : seeds)
            for (const Axis &a : axes)
                for (const json::Value &v : a.values) {
                    combos.push_back({seed, {{a.target, v}}});
                    if (combos.size() > kMaxVariants) break;
                }
    }
    if (combos.empty()) return bail(fail(LENIA_ERR_INVALID_ARG, "the sweep has no variants"));
    if (combos.size() > kMaxVariants)
        return bail(fail(LENIA_ERR_INVALID_ARG, "refusing to run more than %zu variants", kMaxVariants));

    (void)hipSetDevice(ctx->device);
    s->vars.resize(combos.size());
    const size_t px = (size_t)s->base.width * s->base.height;
    int rc = LENIA_OK;
    std::vector<lenia_world *> fftWorlds, directWorlds;
    for (size_t i = 0; i < combos.size() && !rc; ++i) {
        lenia_sweep::Var &v = s->vars[i];
        char id[16];
        std::snprintf(id, sizeof(id), "v%04zu", i);
        v.id = id;
        v.seed = combos[i].first;
        v.overrides = combos[i].second;
        // applyOverrides (main.cpp:296-320): straight into the module's params
        SceneModuleSpec m = *mod;
        bool ok = true;
        for (const auto &ov : v.overrides) {
            const size_t dot = ov.first.find('.');
            if (ov.first.substr(0, dot) != m.name) {
                v.error = "unknown module \"" + ov.first.substr(0, dot) + "\" (target \"" + ov.first + "\")";
                ok = false;
                break;
            }
            set_member(m.params, ov.first.substr(dot + 1), ov.second);
        }
        if (ok && !lenia_setup(m, s->base.width, s->base.height, v.setup, err)) {
            v.error = err;
            ok = false;
        }
        if (ok) {
            const int wrc = lenia_world_create(ctx, &v.setup.desc, &v.world);
            if (wrc == LENIA_ERR_DEVICE || wrc == LENIA_ERR_OUT_OF_MEMORY) {
                rc = wrc;
                break;
            }
            if (wrc) {
                v.error = lenia_last_error();
                ok = false;
            }
        }
        if (ok && v.setup.stamp)
            rc = lenia_world_set_organism(v.world, v.setup.organism.cells.data(), v.setup.organism.width,
                                          v.setup.organism.height, v.setup.stampCount, v.setup.stampRotate ? 1u : 0u);
        if (!rc) rc = dev_alloc(ctx, (void **)&v.layer, px * 4 * sizeof(uint16_t));
        if (!rc) {
            hipError_t e = launch_fill(reinterpret_cast<float *>(v.layer), px * 2, 0.0f, ctx->stream);
            if (e != hipSuccess) rc = hip_fail(e, "clear layer");
        }
        if (rc || !ok) continue;
        // SceneRunner::create resets every module with the scene seed; the
        // static opacity / freeze values decide whether frames run at all
        rc = lenia_world_reset(v.world, v.seed);
        v.visible = module_param(v.setup, {}, "opacity", v.setup.opacity) > 0.02f;
        v.freeze = module_param(v.setup, {}, "freeze", 0.0f) > 0.5f;
        if (!v.visible || v.freeze) continue;
        if (v.setup.desc.convMode == LENIA_CONV_FFT) fftWorlds.push_back(v.world);
        else if (v.setup.desc.convMode == LENIA_CONV_DIRECT) directWorlds.push_back(v.world);
        else s->alone.push_back(i);
    }
    for (auto *ws : {&fftWorlds, &directWorlds}) {
        if (rc || ws->empty()) continue;
        lenia_batch *b = nullptr;
        rc = lenia_batch_create(ctx, ws->data(), (uint32_t)ws->size(), &b);
        if (!rc) s->batches.push_back(b);
    }
    if (!rc) rc = dev_alloc(ctx, (void **)&s->partial, (size_t)luma_stats_blocks(px) * 2 * sizeof(double));
    if (!rc) rc = dev_alloc(ctx, (void **)&s->stats, s->vars.size() * 2 * sizeof(double));
    if (rc) return bail(rc);
    *out = s;
    return LENIA_OK;
}

extern "C" int lenia_sweep_destroy(lenia_sweep *s) {
    if (!s) return LENIA_OK;
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->ctx->stream);
    sweep_release(s);
    delete s;
    return LENIA_OK;
}

extern "C" int lenia_sweep_count(lenia_sweep *s, uint32_t *n) {
    if (!s || !n) return fail(LENIA_ERR_INVALID_ARG, "lenia_sweep_count: NULL argument");
    *n = (uint32_t)s->vars.size();
    return LENIA_OK;
}

extern "C" int lenia_sweep_run(lenia_sweep *s, uint32_t nFrames, uint32_t graphSteps) {
    if (!s) return fail(LENIA_ERR_INVALID_ARG, "lenia_sweep_run: NULL sweep");
    lenia_ctx *ctx = s->ctx;
    (void)hipSetDevice(ctx->device);
    const uint32_t S = std::max(1u, s->substeps);
    // a frozen module encodes its first frame only (SceneRunner.cpp:203-208);
    // frozen variants are few, so they run that frame alone
    for (lenia_sweep::Var &v : s->vars)
        if (v.world && v.visible && v.freeze && nFrames > 0) {
            int rc = lenia_world_frame(v.world, 0, S);
            if (rc) return rc;
        }
    for (lenia_batch *b : s->batches) {
        const uint32_t g = graphSteps && ((uint64_t)nFrames * S) % graphSteps == 0 ? graphSteps : 0u;
        int rc = lenia_batch_frames(b, 0, nFrames, S, g);
        if (rc) return rc;
    }
    for (size_t i : s->alone)
        for (uint32_t f = 0; f < nFrames; ++f) {
            int rc = lenia_world_frame(s->vars[i].world, f, S);
            if (rc) return rc;
        }
    // colour layers and their luma sums, then one read-back
    const size_t px = (size_t)s->base.width * s->base.height;
    hipError_t e = hipSuccess;
    for (size_t i = 0; i < s->vars.size() && e == hipSuccess; ++i) {
        lenia_sweep::Var &v = s->vars[i];
        if (v.world && v.visible) {
            float *state = nullptr;
            int rc = lenia_world_state_device(v.world, &state);
            if (rc) return rc;
            e = launch_colormap(state, (int)v.setup.desc.width, (int)v.setup.desc.height, v.setup.cmap, v.layer,
                                (int)s->base.width, (int)s->base.height, ctx->stream);
        }
        if (e == hipSuccess) e = launch_luma_stats(v.layer, px, s->partial, s->stats + 2 * i, ctx->stream);
    }
    std::vector<double> st(2 * s->vars.size());
    if (e == hipSuccess)
        e = hipMemcpyAsync(st.data(), s->stats, st.size() * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "sweep activity");
    for (size_t i = 0; i < s->vars.size(); ++i) {
        lenia_sweep::Var &v = s->vars[i];
        // computeLumaStats / isAlive (main.cpp:67-90)
        const double mean = st[2 * i] / (double)px;
        double var = st[2 * i + 1] / (double)px - mean * mean;
        if (var < 0.0) var = 0.0;
        v.mean = v.error.empty() ? mean : 0.0;
        v.sd = v.error.empty() ? std::sqrt(var) : 0.0;
        v.alive = v.mean >= 0.01 && v.mean <= 0.7 && v.sd >= 0.02;
    }
    s->framesRun = nFrames;
    return LENIA_OK;
}

extern "C" int lenia_sweep_activity(lenia_sweep *s, double *meanLuma, double *stddevLuma, int *alive, uint32_t n) {
    if (!s) return fail(LENIA_ERR_INVALID_ARG, "lenia_sweep_activity: NULL sweep");
    if (n != s->vars.size()) return fail(LENIA_ERR_INVALID_ARG, "activity: n != variant count");
    for (uint32_t i = 0; i < n; ++i) {
        if (meanLuma) meanLuma[i] = s->vars[i].mean;
        if (stddevLuma) stddevLuma[i] = s->vars[i].sd;
        if (alive) alive[i] = s->vars[i].alive ? 1 : 0;
    }
    return LENIA_OK;
}

extern "C" int lenia_sweep_world(lenia_sweep *s, uint32_t v, lenia_world **world) {
    if (!s || !world) return fail(LENIA_ERR_INVALID_ARG, "lenia_sweep_world: NULL argument");
    if (v >= s->vars.size()) return fail(LENIA_ERR_INVALID_ARG, "variant %u out of range", v);
    *world = s->vars[v].world;
    return LENIA_OK;
}

extern "C" int lenia_sweep_report(lenia_sweep *s, char *buf, size_t cap, size_t *needed) {
    if (!s || !needed) return fail(LENIA_ERR_INVALID_ARG, "lenia_sweep_report: NULL argument");
    auto num = [](double x) {
        json::Value v;
        v.kind = json::Value::Number;
        v.num = x;
        return v;
    };
    auto str = [](const std::string &x) {
        json::Value v;
        v.kind = json::Value::String;
        v.str = x;
        return v;
    };
    json::Value doc;
    doc.kind = json::Value::Object;
    doc.obj.emplace_back("mode", str(s->mode));
    doc.obj.emplace_back("frames", num(s->framesRun));
    doc.obj.emplace_back("width", num(s->base.width));
    doc.obj.emplace_back("height", num(s->base.height));
    doc.obj.emplace_back("substeps", num(std::max(1u, s->substeps)));
    doc.obj.emplace_back("device", str("MI355X (gfx950)"));
    doc.obj.emplace_back("variantCount", num((double)s->vars.size()));
    json::Value list;
    list.kind = json::Value::Array;
    for (const auto &v : s->vars) {
        json::Value e;
        e.kind = json::Value::Object;
        e.obj.emplace_back("id", str(v.id));
        e.obj.emplace_back("seed", num(v.seed));
        json::Value ov;
        ov.kind = json::Value::Object;
        for (const auto &o : v.overrides) ov.obj.emplace_back(o.first, o.second);
        e.obj.emplace_back("overrides", ov);
        if (!v.error.empty()) e.obj.emplace_back("error", str(v.error));
        json::Value act;
        act.kind = json::Value::Object;
        act.obj.emplace_back("meanLuma", num(v.mean));
        act.obj.emplace_back("stddevLuma", num(v.sd));
        json::Value al;
        al.kind = json::Value::Bool;
        al.b = v.alive;
        act.obj.emplace_back("alive", al);
        e.obj.emplace_back("activity", act);
        list.arr.push_back(std::move(e));
    }
    doc.obj.emplace_back("variants", list);
    const std::string text = json::dump(doc);
    *needed = text.size() + 1;
    if (buf) {
        if (cap < text.size() + 1) return fail(LENIA_ERR_INVALID_ARG, "report: buffer too small (%zu bytes needed)", text.size() + 1);
        std::memcpy(buf, text.c_str(), text.size() + 1);
    }
    return LENIA_OK;
}
