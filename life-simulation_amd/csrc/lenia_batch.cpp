// Batched stepping of many independent worlds (the PresetLab sweep, which the
// reference runs one SceneRunner per variant, one after the other:
// Apps/LifePresetLab/main.cpp:718-806).
//
// A batch keeps a device table of world descriptors and one tile list per
// pass spanning every world; a step is one column-pass launch and one
// row-pass launch over all tiles (FFT form) or one launch (direct form), and
// step sequences are replayed from a captured hipGraph.  Each world keeps its
// own buffers, FFT plans and step clock, so a world gives the same bits
// stepped alone or in any batch: the plan of a length never depends on the
// launch shape, and tile shapes only change which threads run a butterfly.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "host_internal.h"
#include "launch.h"

using namespace lenia;

struct lenia_batch {
    lenia_ctx *ctx = nullptr;
    std::vector<lenia_world *> worlds;
    bool isFft = false;
    bool extras = false;
    int rowT = 256, colT = 256;
    bool rowKpar = true, colKpar = true;
    // largest (K) x nt x W (rows) and (K) x nt x H (columns) over the tiles
    int rowCells = 0, colCells = 0;
    void *devWorlds = nullptr;
    int2 *rowTiles = nullptr, *colTiles = nullptr;
    int nRowTiles = 0, nColTiles = 0;
    size_t ldsRows = 0, ldsCols = 0;
    hipGraphExec_t graph = nullptr;
    uint32_t graphSteps = 0;
    std::vector<uint32_t> versions;  // world descriptor versions the table holds
    // launch shapes fixed by lenia_batch_autotune (0: the default choice)
    int tuneRowT = 0, tuneColT = 0;
    bool tuneRowKpar = true, tuneColKpar = true;
    // wave executor: waves per workgroup of each pass (0: every world runs
    // the workgroup-LDS executor).  Worlds whose plan has a wave kernel run
    // it, one launch per values-per-lane class; the others form the LDS
    // group, whose tiles come first in the tile list (nRowTiles / nColTiles
    // of them).
    int rowWaveNW = 0, colWaveNW = 4;
    // the launches of one pass (LDS group, wave classes) run side by side:
    // forked from the context stream onto these streams and joined back
    static constexpr int kSide = 3;
    hipStream_t side[kSide] = {};
    hipEvent_t evFork = nullptr, evJoin[kSide] = {};
    struct WaveGroup {
        int vc = 0, first = 0, count = 0;
        size_t lds = 0;
    };
    std::vector<WaveGroup> rowWaves, colWaves;
};

static void batch_release_tables(lenia_batch *b) {
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    dev_free(b->ctx, b->devWorlds);
    dev_free(b->ctx, b->rowTiles);
    dev_free(b->ctx, b->colTiles);
    b->devWorlds = nullptr;
    b->rowTiles = b->colTiles = nullptr;
    b->ldsRows = b->ldsCols = 0;
}

static void batch_release(lenia_batch *b) {
    batch_release_tables(b);
    for (int i = 0; i < lenia_batch::kSide; ++i) {
        if (b->side[i]) (void)hipStreamDestroy(b->side[i]);
        if (b->evJoin[i]) (void)hipEventDestroy(b->evJoin[i]);
        b->side[i] = nullptr;
        b->evJoin[i] = nullptr;
    }
    if (b->evFork) (void)hipEventDestroy(b->evFork);
    b->evFork = nullptr;
}

// launch(i, stream) for i < n: launch 0 on the context stream, the others on
// side streams that wait for everything before them and that the context
// stream waits for afterwards (under graph capture: parallel branches)
template <class F>
static hipError_t batch_fanout(lenia_batch *b, int n, F launch) {
    hipStream_t s = b->ctx->stream;
    if (n <= 1) return n == 1 ? launch(0, s) : hipSuccess;
    if (n > lenia_batch::kSide + 1) return hipErrorInvalidValue;
    hipError_t e = hipEventRecord(b->evFork, s);
    for (int i = 1; i < n && e == hipSuccess; ++i) e = hipStreamWaitEvent(b->side[i - 1], b->evFork, 0);
    for (int i = 0; i < n && e == hipSuccess; ++i) e = launch(i, i ? b->side[i - 1] : s);
    for (int i = 1; i < n && e == hipSuccess; ++i) {
        e = hipEventRecord(b->evJoin[i - 1], b->side[i - 1]);
        if (e == hipSuccess) e = hipStreamWaitEvent(s, b->evJoin[i - 1], 0);
    }
    return e;
}

static int batch_setup_fft(lenia_batch *b);
static int batch_setup_direct(lenia_batch *b);

// (re)build the world table and tiles from the worlds' current descriptors
static int batch_setup(lenia_batch *b) {
    batch_release_tables(b);
    b->versions.clear();
    for (lenia_world *w : b->worlds) b->versions.push_back(w->version);
    return b->isFft ? batch_setup_fft(b) : batch_setup_direct(b);
}

// transforms per tile for one world at workgroup size T: about 12 values per
// thread, a power of two, and a shape the plan, registers and LDS allow
static int batch_nt(const FftPlan &p, bool rows, int count, int K, int T, bool kpar) {
    int target = std::max(1, (12 * T) / ((kpar ? K : 1) * p.n));
    int nt = 1;
    while (nt * 2 <= target && nt * 2 <= count && nt * 2 <= 64) nt *= 2;
    while (nt > 1 && !shape_ok(p, rows, nt, K, T, kpar)) nt /= 2;
    return shape_ok(p, rows, nt, K, T, kpar) ? nt : 0;
}

// lanes per transform of a world's pass on the wave executor, 0: LDS executor
static int batch_wave_lanes(const lenia_batch *b, const lenia_world *w, bool rows) {
    const int nw = rows ? b->rowWaveNW : b->colWaveNW;
    return nw ? wave_lanes(rows ? w->planW.code : w->planH.code) : 0;
}

static int batch_setup_fft(lenia_batch *b) {
    lenia_ctx *ctx = b->ctx;
    const int n = (int)b->worlds.size();
    // one workgroup size per pass for the LDS group; side-by-side kernels
    // when every world of it allows it (or the tuned shapes)
    const int Ts[4] = {256, 512, 128, 64};
    bool found = false;
    for (int kp = 1; kp >= 0 && !found; --kp)
        for (int T : Ts) {
            if (b->tuneRowT && (T != b->tuneRowT || (kp != 0) != b->tuneRowKpar)) continue;
            bool ok = true;
            for (lenia_world *w : b->worlds) {
                const int K = (int)w->desc.nKernels;
                if (!batch_wave_lanes(b, w, true) && !batch_nt(w->planW, true, w->desc.height / 2, K, T, kp != 0))
                    ok = false;
            }
            if (ok) {
                b->rowT = T;
                b->rowKpar = kp != 0;
                found = true;
                break;
            }
        }
    if (!found) return fail(LENIA_ERR_UNSUPPORTED, "batch: no common row-pass shape");
    found = false;
    for (int kp = 1; kp >= 0 && !found; --kp)
        for (int T : Ts) {
            if (b->tuneColT && (T != b->tuneColT || (kp != 0) != b->tuneColKpar)) continue;
            bool ok = true;
            for (lenia_world *w : b->worlds) {
                const int K = (int)w->desc.nKernels;
                if (!batch_wave_lanes(b, w, false) && !batch_nt(w->planH, false, w->desc.width / 2 + 1, K, T, kp != 0))
                    ok = false;
            }
            if (ok) {
                b->colT = T;
                b->colKpar = kp != 0;
                found = true;
                break;
            }
        }
    if (!found) return fail(LENIA_ERR_UNSUPPORTED, "batch: no common column-pass shape");

    std::vector<FftWorldDev> table(n);
    // tiles: the LDS group first, then one run per wave class
    const int kClasses[3] = {8, 16, 24};
    std::vector<int2> rt, ct, rtw[3], ctw[3];
    size_t rlds[3] = {0, 0, 0}, clds[3] = {0, 0, 0};
    b->rowCells = b->colCells = 0;
    for (int i = 0; i < n; ++i) {
        lenia_world *w = b->worlds[i];
        FftWorldDev f = w->fdev;  // same buffers, plans, clock; batch tile shape
        const int K = (int)w->desc.nKernels;
        const int Lr = batch_wave_lanes(b, w, true), Lc = batch_wave_lanes(b, w, false);
        // kernel-split passes (LENIA_WAVES_PER_KERNEL): one group set per workgroup
        const bool rs = b->rowWaveNW == LENIA_WAVES_PER_KERNEL, cs = b->colWaveNW == LENIA_WAVES_PER_KERNEL;
        f.P = Lr ? (rs ? 1 : b->rowWaveNW) * (64 / Lr) : batch_nt(w->planW, true, f.H / 2, K, b->rowT, b->rowKpar);
        f.C = Lc ? (cs ? 1 : b->colWaveNW) * (64 / Lc) : batch_nt(w->planH, false, f.Wh, K, b->colT, b->colKpar);
        f.rowTiles = (f.H / 2 + f.P - 1) / f.P;
        f.colTiles = (f.Wh + f.C - 1) / f.C;
        f.mP = f.P <= 1 ? 0u : (uint32_t)((((uint64_t)1 << 32) + f.P - 1) / f.P);
        f.rowThreads = Lr ? 64 * (rs ? kMaxKernels : b->rowWaveNW) : b->rowT;
        f.colThreads = Lc ? 64 * (cs ? kMaxKernels : b->colWaveNW) : b->colT;
        table[i] = f;
        if (Lr) {
            const int c = (f.W / Lr) / 8 - 1;  // class 8, 16, 24 -> 0, 1, 2
            for (int t = 0; t < f.rowTiles; ++t) rtw[c].push_back(make_int2(i, t));
            rlds[c] = std::max(rlds[c], rs ? wave_split_lds_bytes(f.W, Lr, true) : wave_lds_bytes(f.W, Lr, b->rowWaveNW));
        } else {
            for (int t = 0; t < f.rowTiles; ++t) rt.push_back(make_int2(i, t));
            b->ldsRows = std::max({b->ldsRows, lds_rows(f.W, f.P, K, b->rowKpar, b->rowT),
                                   lds_rows(f.W, f.P, 1, true, b->rowT)});
            b->rowCells = std::max(b->rowCells, (b->rowKpar ? K : 1) * f.P * f.W);
        }
        if (Lc) {
            const int c = (f.H / Lc) / 8 - 1;
            for (int t = 0; t < f.colTiles; ++t) ctw[c].push_back(make_int2(i, t));
            clds[c] = std::max(clds[c], cs ? wave_split_lds_bytes(f.H, Lc, false) : wave_lds_bytes(f.H, Lc, b->colWaveNW));
        } else {
            for (int t = 0; t < f.colTiles; ++t) ct.push_back(make_int2(i, t));
            b->ldsCols = std::max(b->ldsCols, lds_cols(f.H, f.C, K, b->colKpar));
            b->colCells = std::max(b->colCells, (b->colKpar ? K : 1) * f.C * f.H);
        }
    }
    b->nRowTiles = (int)rt.size();
    b->nColTiles = (int)ct.size();
    b->rowWaves.clear();
    b->colWaves.clear();
    for (int c = 0; c < 3; ++c) {
        if (!rtw[c].empty()) {
            lenia_batch::WaveGroup gr;
            gr.vc = kClasses[c];
            gr.first = (int)rt.size();
            gr.count = (int)rtw[c].size();
            gr.lds = rlds[c];
            b->rowWaves.push_back(gr);
            rt.insert(rt.end(), rtw[c].begin(), rtw[c].end());
        }
        if (!ctw[c].empty()) {
            lenia_batch::WaveGroup gr;
            gr.vc = kClasses[c];
            gr.first = (int)ct.size();
            gr.count = (int)ctw[c].size();
            gr.lds = clds[c];
            b->colWaves.push_back(gr);
            ct.insert(ct.end(), ctw[c].begin(), ctw[c].end());
        }
    }
    int rc = dev_alloc(ctx, &b->devWorlds, sizeof(FftWorldDev) * n);
    if (!rc) rc = dev_alloc(ctx, (void **)&b->rowTiles, sizeof(int2) * rt.size());
    if (!rc) rc = dev_alloc(ctx, (void **)&b->colTiles, sizeof(int2) * ct.size());
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(b->devWorlds, table.data(), sizeof(FftWorldDev) * n,
                           hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->rowTiles, rt.data(), sizeof(int2) * rt.size(), hipMemcpyHostToDevice,
                           ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->colTiles, ct.data(), sizeof(int2) * ct.size(), hipMemcpyHostToDevice,
                           ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? LENIA_OK : hip_fail(e, "batch upload");
}

static int batch_setup_direct(lenia_batch *b) {
    lenia_ctx *ctx = b->ctx;
    const int n = (int)b->worlds.size();
    std::vector<DirectWorldDev> table(n);
    std::vector<int2> rt;
    for (int i = 0; i < n; ++i) {
        lenia_world *w = b->worlds[i];
        // the batch tiles every world 64 x 32 (a world alone may use 64 x 16
        // tiles; each output's arithmetic is the same in both)
        DirectWorldDev d = w->ddev;
        d.tilesY = (d.H + kDirectTileH - 1) / kDirectTileH;
        d.tileH = kDirectTileH;
        table[i] = d;
        for (int t = 0; t < d.tilesX * d.tilesY; ++t) rt.push_back(make_int2(i, t));
    }
    // the split form, for batches of at most one tile per CU, holds the
    // second half's sums in LDS as well
    for (int i = 0; i < n; ++i)
        b->ldsRows = std::max(b->ldsRows, direct_lds_bytes(table[i].half, kDirectTileH,
                                                           direct_split((int)rt.size(), ctx->numCUs)));
    b->nRowTiles = (int)rt.size();
    int rc = dev_alloc(ctx, &b->devWorlds, sizeof(DirectWorldDev) * n);
    if (!rc) rc = dev_alloc(ctx, (void **)&b->rowTiles, sizeof(int2) * rt.size());
    if (rc) return rc;
    hipError_t e = hipMemcpyAsync(b->devWorlds, table.data(), sizeof(DirectWorldDev) * n,
                                  hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(b->rowTiles, rt.data(), sizeof(int2) * rt.size(), hipMemcpyHostToDevice,
                           ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    return e == hipSuccess ? LENIA_OK : hip_fail(e, "batch upload");
}

extern "C" int lenia_batch_create(lenia_ctx *ctx, lenia_world *const *worlds, uint32_t n,
                                  lenia_batch **out) {
    if (!ctx || !worlds || !out || n == 0)
        return fail(LENIA_ERR_INVALID_ARG, "lenia_batch_create: bad argument");
    *out = nullptr;
    (void)hipSetDevice(ctx->device);
    for (uint32_t i = 0; i < n; ++i) {
        if (!worlds[i]) return fail(LENIA_ERR_INVALID_ARG, "lenia_batch_create: world %u is NULL", i);
        if (worlds[i]->ctx != ctx)
            return fail(LENIA_ERR_INVALID_ARG, "lenia_batch_create: world %u has another context", i);
        if (worlds[i]->desc.convMode != worlds[0]->desc.convMode)
            return fail(LENIA_ERR_INVALID_ARG, "lenia_batch_create: all worlds need one convMode");
        for (uint32_t j = 0; j < i; ++j)
            if (worlds[j] == worlds[i])
                return fail(LENIA_ERR_INVALID_ARG, "lenia_batch_create: world %u listed twice", i);
    }
    if (worlds[0]->isMfma)
        return fail(LENIA_ERR_UNSUPPORTED, "lenia_batch_create: MFMA forms are stepped alone");
    lenia_batch *b = new lenia_batch();
    b->ctx = ctx;
    hipError_t es = hipEventCreateWithFlags(&b->evFork, hipEventDisableTiming);
    for (int i = 0; i < lenia_batch::kSide && es == hipSuccess; ++i) {
        es = hipStreamCreateWithFlags(&b->side[i], hipStreamNonBlocking);
        if (es == hipSuccess) es = hipEventCreateWithFlags(&b->evJoin[i], hipEventDisableTiming);
    }
    if (es != hipSuccess) {
        batch_release(b);
        delete b;
        return hip_fail(es, "batch streams");
    }
    b->worlds.assign(worlds, worlds + n);
    b->isFft = worlds[0]->isFft;
    for (lenia_world *w : b->worlds) b->extras = b->extras || w->extras();
    int rc = batch_setup(b);
    if (rc) {
        batch_release(b);
        delete b;
        return rc;
    }
    *out = b;
    return LENIA_OK;
}

extern "C" int lenia_batch_destroy(lenia_batch *b) {
    if (!b) return LENIA_OK;
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    batch_release(b);
    delete b;
    return LENIA_OK;
}

// one step of every world, `off` steps after its stored clock; with `ev`,
// pass p (columns, rows) records ev[2p] and ev[2p+1] around its launches
static hipError_t batch_enqueue_step(lenia_batch *b, const hipEvent_t *ev, uint32_t off) {
    lenia_ctx *ctx = b->ctx;
    hipStream_t s = ctx->stream;
    const FftWorldDev *table = (const FftWorldDev *)b->devWorlds;
    hipError_t e = hipSuccess;
    if (b->isFft) {
        // launch 0: the LDS group when it has tiles; then the wave classes
        const int cl = b->nColTiles ? 1 : 0, rl = b->nRowTiles ? 1 : 0;
        if (ev) e = hipEventRecord(ev[0], s);
        if (e == hipSuccess)
            e = batch_fanout(b, cl + (int)b->colWaves.size(), [&](int i, hipStream_t st) {
                if (i < cl)
                    return launch_fft_cols(b->colT, b->colKpar, table, b->colTiles, b->nColTiles,
                                           b->nColTiles, b->ldsCols, st, b->colCells);
                const auto &gr = b->colWaves[i - cl];
                if (b->colWaveNW == LENIA_WAVES_PER_KERNEL)
                    return launch_wave_cols_split(gr.vc, table, b->colTiles + gr.first, gr.count, gr.lds, st);
                return launch_wave_cols(b->colWaveNW, gr.vc, table, b->colTiles + gr.first, gr.count,
                                        gr.lds, st);
            });
        if (ev && e == hipSuccess) e = hipEventRecord(ev[1], s);
        if (ev && e == hipSuccess) e = hipEventRecord(ev[2], s);
        if (e == hipSuccess)
            e = batch_fanout(b, rl + (int)b->rowWaves.size(), [&](int i, hipStream_t st) {
                if (i < rl)
                    return launch_fft_rows(1, b->rowT, b->extras, b->rowKpar, table, b->rowTiles,
                                           b->nRowTiles, b->nRowTiles, b->ldsRows, off, 0, nullptr, st,
                                           b->rowCells);
                const auto &gr = b->rowWaves[i - rl];
                if (b->rowWaveNW == LENIA_WAVES_PER_KERNEL)
                    return launch_wave_rows_split(b->extras, gr.vc, table, b->rowTiles + gr.first, gr.count,
                                                  gr.lds, off, st);
                return launch_wave_rows(1, b->rowWaveNW, gr.vc, b->extras, table, b->rowTiles + gr.first,
                                        gr.count, gr.lds, off, 0, nullptr, st);
            });
        if (ev && e == hipSuccess) e = hipEventRecord(ev[3], s);
    } else {
        if (ev) launch_events() = {ev[0], ev[1]};
        e = launch_direct(0, (const DirectWorldDev *)b->devWorlds, b->rowTiles, b->nRowTiles,
                          b->nRowTiles, b->ldsRows, off, 0, nullptr, s, b->ctx->numCUs, kDirectTileH);
        launch_events() = LaunchEvents();
    }
    return e;
}

// every world's stored clock moves `by` steps on
static hipError_t batch_advance(lenia_batch *b, uint32_t by) {
    const int n = (int)b->worlds.size();
    if (b->isFft) return launch_advance_clocks((const FftWorldDev *)b->devWorlds, n, by, b->ctx->stream);
    return launch_advance_clocks((const DirectWorldDev *)b->devWorlds, n, by, b->ctx->stream);
}

// n steps of every world (offsets 0 .. n-1), then the clocks move n steps on
static hipError_t batch_enqueue_steps(lenia_batch *b, uint32_t n) {
    hipError_t e = hipSuccess;
    for (uint32_t i = 0; i < n && e == hipSuccess; ++i) e = batch_enqueue_step(b, nullptr, i);
    if (e == hipSuccess) e = batch_advance(b, n);
    return e;
}

// table up to date, every world initialised and holding its row spectrum
static int batch_prepare(lenia_batch *b) {
    lenia_ctx *ctx = b->ctx;
    bool stale = false;
    for (size_t i = 0; i < b->worlds.size(); ++i) stale = stale || b->versions[i] != b->worlds[i]->version;
    if (stale) {
        // a world's parameters changed since the table was built
        b->extras = false;
        for (lenia_world *w : b->worlds) b->extras = b->extras || w->extras();
        int rc = batch_setup(b);
        if (rc) return rc;
    }
    for (lenia_world *w : b->worlds) {
        int rc = LENIA_OK;
        if (w->needsInit) {
            rc = world_apply_init(w);
            hipError_t e = rc ? hipSuccess
                              : launch_set_clock(w->clock, w->frameNext, 0, 1, w->parity(), ctx->stream);
            if (e != hipSuccess) rc = hip_fail(e, "set clock");
        }
        if (!rc) rc = world_ensure_spectrum(w);
        if (rc) return rc;
    }
    return LENIA_OK;
}

// Empirical launch shapes for the whole batch (FFT form): each pass's
// workgroup size and side-by-side choice are timed over nSteps steps with
// per-kernel timestamps; the worlds' states and clocks are restored.  Plans
// do not change, so results stay bit-identical to stepping each world alone.
static int batch_run_steps(lenia_batch *b, uint32_t nSteps, uint32_t graphSteps);

extern "C" int lenia_batch_autotune(lenia_batch *b, uint32_t nSteps, double *msPerStep) {
    if (!b) return fail(LENIA_ERR_INVALID_ARG, "lenia_batch_autotune: NULL batch");
    if (!b->isFft) {
        if (msPerStep) *msPerStep = 0.0;
        return LENIA_OK;
    }
    lenia_ctx *ctx = b->ctx;
    (void)hipSetDevice(ctx->device);
    nSteps = std::max(nSteps, 4u);
    int rc = batch_prepare(b);
    if (rc) return rc;
    // save states and clocks
    size_t total = 0;
    for (lenia_world *w : b->worlds) total += (size_t)w->desc.width * w->desc.height;
    float *saved = nullptr;
    rc = dev_alloc(ctx, (void **)&saved, total * sizeof(float));
    if (rc) return rc;
    std::vector<StepClock> clocks(b->worlds.size());
    hipError_t e = hipSuccess;
    size_t off = 0;
    for (size_t i = 0; i < b->worlds.size() && e == hipSuccess; ++i) {
        lenia_world *w = b->worlds[i];
        const size_t cells = (size_t)w->desc.width * w->desc.height;
        e = hipMemcpyAsync(saved + off, w->current_state(), cells * sizeof(float),
                           hipMemcpyDeviceToDevice, ctx->stream);
        if (e == hipSuccess)
            e = hipMemcpyAsync(&clocks[i], w->clock, sizeof(StepClock), hipMemcpyDeviceToHost, ctx->stream);
        off += cells;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = hip_fail(e, "batch autotune save");
    if (b->graph) (void)hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    auto time_now = [&](double ms[2]) -> int {
        int r = batch_setup(b);
        if (r) return r;
        hipError_t et = batch_enqueue_steps(b, 2);
        if (et == hipSuccess)
            et = time_steps(ctx, nSteps, 2, [&](const hipEvent_t *ev) {
                hipError_t es = batch_enqueue_step(b, ev, 0);
                return es == hipSuccess ? batch_advance(b, 1) : es;
            }, ms);
        return et == hipSuccess ? LENIA_OK : hip_fail(et, "batch autotune timing");
    };
    double best[2] = {1e30, 1e30};
    int bestRowT = 0, bestColT = 0;
    bool bestRowKp = true, bestColKp = true;
    // columns keep their default while the row shapes are timed, and back
    b->tuneColT = 0;
    // the workgroup-LDS executor for every world first, the wave executor last
    b->rowWaveNW = b->colWaveNW = 0;
    for (int kp = 1; kp >= 0 && !rc; --kp)
        for (int T = 64; T <= 512 && !rc; T *= 2) {
            b->tuneRowT = T;
            b->tuneRowKpar = kp != 0;
            double ms[2];
            if (time_now(ms) != LENIA_OK) continue;  // shape not valid for every world
            if (ms[1] < best[1]) {
                best[1] = ms[1];
                bestRowT = T;
                bestRowKp = kp != 0;
            }
        }
    b->tuneRowT = bestRowT;
    b->tuneRowKpar = bestRowKp;
    for (int kp = 1; kp >= 0 && !rc && bestRowT; --kp)
        for (int T = 64; T <= 512; T *= 2) {
            b->tuneColT = T;
            b->tuneColKpar = kp != 0;
            double ms[2];
            if (time_now(ms) != LENIA_OK) continue;
            if (ms[0] < best[0]) {
                best[0] = ms[0];
                bestColT = T;
                bestColKp = kp != 0;
            }
        }
    b->tuneColT = bestColT;
    b->tuneColKpar = bestColKp;
    // then the wave executor (worlds whose plan has a wave kernel) at 1, 2
    // and 4 waves per workgroup and kernel-split, against the best so far
    int bestRowNW = 0, bestColNW = 0;
    const int kNw[4] = {1, 2, 4, LENIA_WAVES_PER_KERNEL};
    for (int ni = 0; ni < 4 && !rc && bestRowT && bestColT; ++ni) {
        const int nw = kNw[ni];
        b->rowWaveNW = b->colWaveNW = nw;
        double ms[2];
        if (time_now(ms) != LENIA_OK) continue;
        if (ms[0] < best[0]) {
            best[0] = ms[0];
            bestColNW = nw;
        }
        if (ms[1] < best[1]) {
            best[1] = ms[1];
            bestRowNW = nw;
        }
    }
    // The per-pass times above come from launches timed one by one; the
    // batch runs graph-replayed, where the side-by-side launches of the
    // executors overlap differently.  The finalists are timed that way and
    // the fastest whole step is kept.
    double graphStepMs = 0.0;
    if (!rc && bestRowT && bestColT) {
        std::vector<uint64_t> hostSteps;
        for (lenia_world *w : b->worlds) hostSteps.push_back(w->hostSteps);
        // graph-replayed time per step of the batch as currently configured
        // (negative: the configuration does not run)
        auto graph_step = [&]() -> double {
            constexpr uint32_t kG = 8, kRuns = 4;
            if (batch_setup(b) != LENIA_OK) return -1.0;
            if (batch_run_steps(b, kG, kG) != LENIA_OK) return -1.0;  // capture, warm-up
            if (batch_run_steps(b, kRuns * kG, kG) != LENIA_OK) return -1.0;
            float ms = 0.f;
            if (hipEventSynchronize(ctx->ev1) != hipSuccess ||
                hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1) != hipSuccess)
                return -1.0;
            return ms / (kRuns * kG);
        };
        std::vector<std::pair<int, int>> cand = {{bestRowNW, bestColNW}, {0, 0}};
        for (int ni = 0; ni < 4; ++ni) cand.push_back({bestRowNW, kNw[ni]});
        double bestStep = 1e30;
        std::pair<int, int> pick = cand[0];
        for (size_t i = 0; i < cand.size(); ++i) {
            if (std::find(cand.begin(), cand.begin() + i, cand[i]) != cand.begin() + i) continue;
            b->rowWaveNW = cand[i].first;
            b->colWaveNW = cand[i].second;
            const double per = graph_step();
            if (per > 0.0 && per < bestStep) {
                bestStep = per;
                pick = cand[i];
            }
        }
        bestRowNW = pick.first;
        bestColNW = pick.second;
        for (size_t i = 0; i < b->worlds.size(); ++i) b->worlds[i]->hostSteps = hostSteps[i];
        if (bestStep < 1e30) graphStepMs = bestStep;
    }
    b->rowWaveNW = bestRowNW;
    b->colWaveNW = bestColNW;
    const int rc2 = batch_setup(b);
    // restore
    off = 0;
    e = hipSuccess;
    for (size_t i = 0; i < b->worlds.size() && e == hipSuccess; ++i) {
        lenia_world *w = b->worlds[i];
        const size_t cells = (size_t)w->desc.width * w->desc.height;
        e = hipMemcpyAsync(w->current_state(), saved + off, cells * sizeof(float),
                           hipMemcpyDeviceToDevice, ctx->stream);
        if (e == hipSuccess)
            e = launch_set_clock(w->clock, clocks[i].frame, clocks[i].sub, clocks[i].substeps,
                                 clocks[i].parity, ctx->stream);
        w->specValid = false;
        off += cells;
    }
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    dev_free(ctx, saved);
    if (rc) return rc;
    if (!bestRowT || !bestColT) return fail(LENIA_ERR_UNSUPPORTED, "batch autotune: no valid shape");
    if (rc2) return rc2;
    if (e != hipSuccess) return hip_fail(e, "batch autotune restore");
    if (msPerStep) *msPerStep = graphStepMs > 0.0 ? graphStepMs : best[0] + best[1];
    return LENIA_OK;
}

// The tile-queue grids (a resident grid taking tiles from per-XCD queues) are
// withdrawn: they ran 3x slower than one workgroup per tile on the 256-world
// sweep (0.57 against 0.19 ms per step) and the autotuner never chose them.
// The calls remain for callers built against them: 0, 0 is accepted (one
// workgroup per tile, the only launch form), anything else is refused.
extern "C" int lenia_batch_set_grids(lenia_batch *b, uint32_t colPerCU, uint32_t rowPerCU) {
    if (!b) return fail(LENIA_ERR_INVALID_ARG, "lenia_batch_set_grids: NULL batch");
    if (colPerCU || rowPerCU)
        return fail(LENIA_ERR_UNSUPPORTED,
                    "lenia_batch_set_grids: tile-queue grids are withdrawn (one workgroup per tile only)");
    return LENIA_OK;
}

extern "C" int lenia_batch_set_waves(lenia_batch *b, uint32_t rowWaves, uint32_t colWaves) {
    if (!b) return fail(LENIA_ERR_INVALID_ARG, "lenia_batch_set_waves: NULL batch");
    auto ok = [](uint32_t v) { return v <= 4; };
    if (!ok(rowWaves) || !ok(colWaves))
        return fail(LENIA_ERR_INVALID_ARG,
                    "lenia_batch_set_waves: %u / %u (0, 1, 2, 4 waves per workgroup or 3, kernel-split)",
                    rowWaves, colWaves);
    if (!b->isFft && (rowWaves || colWaves))
        return fail(LENIA_ERR_UNSUPPORTED, "lenia_batch_set_waves: the wave executor needs the FFT form");
    (void)hipSetDevice(b->ctx->device);
    (void)hipStreamSynchronize(b->ctx->stream);
    b->rowWaveNW = (int)rowWaves;
    b->colWaveNW = (int)colWaves;
    if (b->isFft) {
        const int rc = batch_setup(b);
        if (rc) return rc;
    }
    return LENIA_OK;
}

extern "C" int lenia_batch_get_waves(const lenia_batch *b, uint32_t *rowWaves, uint32_t *colWaves) {
    if (!b || !rowWaves || !colWaves) return fail(LENIA_ERR_INVALID_ARG, "lenia_batch_get_waves: bad argument");
    *rowWaves = (uint32_t)b->rowWaveNW;
    *colWaves = (uint32_t)b->colWaveNW;
    return LENIA_OK;
}

extern "C" int lenia_batch_get_grids(const lenia_batch *b, uint32_t *colPerCU, uint32_t *rowPerCU) {
    if (!b || !colPerCU || !rowPerCU) return fail(LENIA_ERR_INVALID_ARG, "lenia_batch_get_grids: bad argument");
    *colPerCU = *rowPerCU = 0;
    return LENIA_OK;
}

// nSteps steps of every world after batch_prepare (graph replays when
// graphSteps > 0), bracketed by the context's timing events
static int batch_run_steps(lenia_batch *b, uint32_t nSteps, uint32_t graphSteps) {
    lenia_ctx *ctx = b->ctx;
    (void)hipEventRecord(ctx->ev0, ctx->stream);
    if (graphSteps) {
        if (!b->graph || b->graphSteps != graphSteps) {
            if (b->graph) (void)hipGraphExecDestroy(b->graph);
            b->graph = nullptr;
            hipGraph_t g = nullptr;
            hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) return hip_fail(e, "batch graph capture");
            hipError_t el = batch_enqueue_steps(b, graphSteps);
            e = hipStreamEndCapture(ctx->stream, &g);
            if (el != hipSuccess) e = el;
            if (e == hipSuccess) e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
            if (g) (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                b->graph = nullptr;
                return hip_fail(e, "batch graph instantiate");
            }
            b->graphSteps = graphSteps;
        }
        for (uint32_t r = 0; r < nSteps / graphSteps; ++r) {
            hipError_t e = hipGraphLaunch(b->graph, ctx->stream);
            if (e != hipSuccess) return hip_fail(e, "batch graph launch");
        }
    } else {
        hipError_t e = batch_enqueue_steps(b, nSteps);
        if (e != hipSuccess) return hip_fail(e, "batch step launch");
    }
    (void)hipEventRecord(ctx->ev1, ctx->stream);
    ctx->timed = true;
    ctx->last0 = ctx->ev0;
    ctx->last1 = ctx->ev1;
    for (lenia_world *w : b->worlds) w->hostSteps += nSteps;
    return LENIA_OK;
}

extern "C" int lenia_batch_step(lenia_batch *b, uint32_t nSteps, uint32_t graphSteps) {
    if (!b) return fail(LENIA_ERR_INVALID_ARG, "lenia_batch_step: NULL batch");
    if (graphSteps && nSteps % graphSteps)
        return fail(LENIA_ERR_INVALID_ARG, "lenia_batch_step: nSteps %u is not a multiple of %u",
                    nSteps, graphSteps);
    (void)hipSetDevice(b->ctx->device);
    int rc = batch_prepare(b);
    if (rc) return rc;
    rc = batch_run_steps(b, nSteps, graphSteps);
    if (rc) return rc;
    for (lenia_world *w : b->worlds) w->frameNext += nSteps;
    return LENIA_OK;
}

extern "C" int lenia_batch_frames(lenia_batch *b, uint32_t frame0, uint32_t nFrames, uint32_t substeps,
                                  uint32_t graphSteps) {
    if (!b) return fail(LENIA_ERR_INVALID_ARG, "lenia_batch_frames: NULL batch");
    const uint32_t S = std::max(1u, substeps);
    const uint64_t total = (uint64_t)nFrames * S;
    if (total > 0xFFFFFFFFull) return fail(LENIA_ERR_INVALID_ARG, "lenia_batch_frames: too many steps");
    const uint32_t nSteps = (uint32_t)total;
    if (graphSteps && nSteps % graphSteps)
        return fail(LENIA_ERR_INVALID_ARG, "lenia_batch_frames: %u steps are not a multiple of %u",
                    nSteps, graphSteps);
    (void)hipSetDevice(b->ctx->device);
    int rc = batch_prepare(b);
    if (rc) return rc;
    // every world's clock at (frame0, sub 0, S substeps); the step offsets
    // then walk the subs of consecutive frames
    const int n = (int)b->worlds.size();
    hipError_t e = b->isFft ? launch_frame_clocks((const FftWorldDev *)b->devWorlds, n, frame0, S, b->ctx->stream)
                            : launch_frame_clocks((const DirectWorldDev *)b->devWorlds, n, frame0, S, b->ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "frame clocks");
    rc = batch_run_steps(b, nSteps, graphSteps);
    if (rc) return rc;
    for (lenia_world *w : b->worlds) w->frameNext = frame0 + nFrames;
    return LENIA_OK;
}
