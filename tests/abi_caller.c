/* A plain C caller of the Lenia C ABI, built by tests/test_abi_cpu.py with gcc
 * against include/lenia_mi355x.h and linked to liblenia_mi355x.so.  It prints
 * the header's struct layouts (sizeof / offsetof, for comparison with the
 * Python mirror), then drives the calls that need no GPU: NULL arguments,
 * lenia_last_error, lenia_version, the colour-map defaults and
 * lenia_scene_describe on the scene given as argv[1]. */
#include <stddef.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lenia_mi355x.h"

#define SZ(T) printf("size %s %zu\n", #T, sizeof(T))
#define OFF(T, f) printf("offset %s %s %zu\n", #T, #f, offsetof(T, f))

static void layouts(void) {
    SZ(lenia_kernel_desc);
    OFF(lenia_kernel_desc, shape);
    OFF(lenia_kernel_desc, core);
    OFF(lenia_kernel_desc, radius);
    OFF(lenia_kernel_desc, shellMu);
    OFF(lenia_kernel_desc, shellSigma);
    OFF(lenia_kernel_desc, nBetas);
    OFF(lenia_kernel_desc, betas);
    OFF(lenia_kernel_desc, mu);
    OFF(lenia_kernel_desc, sigma);
    OFF(lenia_kernel_desc, weight);
    SZ(lenia_world_desc);
    OFF(lenia_world_desc, width);
    OFF(lenia_world_desc, height);
    OFF(lenia_world_desc, convMode);
    OFF(lenia_world_desc, growthMode);
    OFF(lenia_world_desc, dt);
    OFF(lenia_world_desc, noiseAmount);
    OFF(lenia_world_desc, muJitter);
    OFF(lenia_world_desc, injectAmount);
    OFF(lenia_world_desc, injectCount);
    OFF(lenia_world_desc, radius);
    OFF(lenia_world_desc, nKernels);
    OFF(lenia_world_desc, kernels);
    OFF(lenia_world_desc, initMode);
    OFF(lenia_world_desc, initX0);
    OFF(lenia_world_desc, initY0);
    OFF(lenia_world_desc, initW);
    OFF(lenia_world_desc, initH);
    OFF(lenia_world_desc, initCoverage);
    OFF(lenia_world_desc, initScale);
    OFF(lenia_world_desc, fftPlanW);
    OFF(lenia_world_desc, fftPlanH);
    OFF(lenia_world_desc, rowPairsPerTile);
    OFF(lenia_world_desc, colsPerTile);
    OFF(lenia_world_desc, rowThreads);
    OFF(lenia_world_desc, colThreads);
    OFF(lenia_world_desc, rowEngine);
    OFF(lenia_world_desc, colEngine);
    SZ(lenia_param_grads);
    OFF(lenia_param_grads, dMu);
    OFF(lenia_param_grads, dSigma);
    OFF(lenia_param_grads, dWeight);
    OFF(lenia_param_grads, dBetas);
    OFF(lenia_param_grads, dDt);
    SZ(lenia_colormap);
    OFF(lenia_colormap, a);
    OFF(lenia_colormap, b);
    OFF(lenia_colormap, c);
    OFF(lenia_colormap, d);
    OFF(lenia_colormap, inputScale);
    OFF(lenia_colormap, inputBias);
    OFF(lenia_colormap, exposure);
    OFF(lenia_colormap, channel);
    SZ(lenia_scene_info);
    OFF(lenia_scene_info, width);
    OFF(lenia_scene_info, height);
    OFF(lenia_scene_info, seed);
    OFF(lenia_scene_info, substeps);
    OFF(lenia_scene_info, frameIndex);
    OFF(lenia_scene_info, nModules);
}

#define RC(call)                                              \
    do {                                                      \
        int rc_ = (int)(call);                                \
        printf("rc %s %d %s\n", #call, rc_, lenia_last_error()); \
    } while (0)

static void null_paths(void) {
    lenia_world_desc d;
    memset(&d, 0, sizeof d);
    lenia_param_grads g;
    double ms = 0.0;
    RC(lenia_ctx_create(0, NULL));
    RC(lenia_ctx_sync(NULL));
    RC(lenia_ctx_last_gpu_ms(NULL, &ms));
    RC(lenia_world_create(NULL, &d, NULL));
    RC(lenia_world_step(NULL, 1));
    RC(lenia_world_frame(NULL, 0, 1));
    RC(lenia_world_read_state(NULL, NULL, 0));
    RC(lenia_world_get_desc(NULL, &d));
    RC(lenia_rollout_backward(NULL, 4, 0, NULL, NULL, NULL, &g, 0));
    RC(lenia_batch_step(NULL, 1, 0));
    RC(lenia_scene_describe(NULL, 0, 0, &d, NULL, NULL, 0));
}

int main(int argc, char **argv) {
    layouts();
    printf("version %s\n", lenia_version());
    null_paths();
    lenia_colormap cm;
    lenia_colormap_lenia_defaults(&cm);
    printf("palette d %.2f %.2f %.2f channel %u\n", cm.d[0], cm.d[1], cm.d[2], cm.channel);
    if (argc > 1) {
        FILE *f = fopen(argv[1], "rb");
        if (!f) return 2;
        fseek(f, 0, SEEK_END);
        long n = ftell(f);
        fseek(f, 0, SEEK_SET);
        char *text = (char *)malloc((size_t)n);
        if (!text || fread(text, 1, (size_t)n, f) != (size_t)n) return 2;
        fclose(f);
        lenia_world_desc d;
        lenia_colormap c;
        char name[64];
        int rc = lenia_scene_describe(text, (size_t)n, 0, &d, &c, name, sizeof name);
        printf("describe rc %d name %s size %ux%u conv %u growth %u K %u R %u dt %.3f\n", rc, name,
               d.width, d.height, d.convMode, d.growthMode, d.nKernels, d.radius, d.dt);
        for (uint32_t k = 0; k < d.nKernels && rc == 0; ++k)
            printf("kernel %u radius %.3f mu %.3f sigma %.3f weight %.3f nBetas %u\n", k,
                   d.kernels[k].radius, d.kernels[k].mu, d.kernels[k].sigma, d.kernels[k].weight,
                   d.kernels[k].nBetas);
        free(text);
    }
    return 0;
}
