/*
 * lenia_oracle.c -- FP64 CPU restatement of the LifeSimSuite Lenia hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, never part of
 * the product path.  Only tests/, the smoke entry point and the benchmark's
 * CPU-baseline leg load it.
 *
 * Reference files restated (paths relative to the reference repository):
 *   Shaders/Common/Random.metal:5-19      pcg_hash, rand01
 *   LifeCore/Math/Random.h:11-44          SplitMix64, pcgHash, deriveSeed
 *   Shaders/Common/CommonTypes.metal:39-45 wrapCoord
 *   Modules/FieldModules/Lenia.cpp:95-125  buildRingKernelImage
 *   Modules/FieldModules/Lenia.cpp:410-436 buildKernelTexture
 *   Modules/FieldModules/Lenia.cpp:448-451, 562-564  seed / step-seed semantics
 *   Shaders/Field/Lenia.metal:32-61        leniaValueNoise / leniaInit
 *   Shaders/Field/Lenia.metal:63-128       leniaStep (direct potential + growth)
 *   Shaders/Field/Lenia.metal:156-232      leniaBell / leniaGrowthMulti
 *   Shaders/Field/Lenia.metal:242-303      leniaStampInit (organism stamps)
 *   Shaders/Common/FFT.metal:13-45         FFT sign/scale convention
 *   Shaders/Render/ColorMap.metal:17-75    colorMapField / colorMapFieldScaled
 *
 * Pinning (each check is independent of the formula it checks):
 *   hashes      known answers computed from an independent transcription, and
 *               equality with the reference's own Random.h compiled into
 *               oracle/_ref (tests/test_oracle.py)
 *   kernels     sum 1; 8-fold symmetry; zero at r=0 and r>R; peak position of
 *               the shell; ring == window for betas=[1], shellMu=0.5; golden
 *               images from an independent numpy builder
 *   potentials  delta field -> kernel; uniform field -> constant; numpy.fft
 *               circular convolution golden vectors; direct == FFT
 *   FFT         numpy.fft golden spectra; delta/constant/cosine closed forms;
 *               Parseval; round trip
 *   growth      uniform-field closed forms, mode 0 A' = clip(c + dt(2 bell(c)-1))
 *               and mode 1 A' = clip(c + dt(sum w bell(c) / sum w - c)) (A' = c
 *               when sum w = 0); zero field stays zero; numpy golden steps
 *   gradient    torch float64 autograd golden vectors; FP64 central finite
 *               differences of lo_rollout (small worlds) and, at the 256x256
 *               T=64 size, of lo_rollout_piece (clip states of the recorded
 *               trajectory held fixed, so no difference straddles a clip kink)
 *   piece       lo_rollout_piece records the same trajectory as lo_rollout
 *               and replaying its own record reproduces it bit for bit
 *   stamps      one stamp is the organism whole in one of the 8 dihedral
 *               orientations (all 8 occur over seeds); unrotated it is the
 *               organism as stored; overlapping stamps take the maximum
 *   colour map  palette closed forms (t = 0, the mono ramp, clamps, bias /
 *               scale / exposure, the channel rule); 2x downsampling = mean of
 *               four texels, wrap at the border, constants stay constant
 *   value noise parity unpinned against the reference (needs Metal FP32);
 *               pinned only by range and lattice-continuity invariants
 */
#include "lenia_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* hashes: Random.metal:5-19, Random.h:34-44                           */
/* ------------------------------------------------------------------ */
uint32_t lo_pcg_hash(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}

uint32_t lo_derive_seed(uint32_t globalSeed, uint32_t tag) {
    return lo_pcg_hash(globalSeed ^ lo_pcg_hash(tag));
}

/* rand01's integer part: pcg(x ^ pcg(y ^ pcg(tag ^ seed))) (Random.metal:16-19) */
uint32_t lo_hash3(uint32_t x, uint32_t y, uint32_t tag, uint32_t seed) {
    return lo_pcg_hash(x ^ lo_pcg_hash(y ^ lo_pcg_hash(tag ^ seed)));
}

double lo_rand01(uint32_t x, uint32_t y, uint32_t tag, uint32_t seed) {
    return (double)(lo_hash3(x, y, tag, seed) >> 8) * (1.0 / 16777216.0);
}

/* LeniaModule::reset (Lenia.cpp:448-451): seed_ = deriveSeed(sceneSeed, 'LEN1') */
uint32_t lo_module_seed(uint32_t sceneSeed) { return lo_derive_seed(sceneSeed, 0x4C454E31u); }

/* per-substep seed (Lenia.cpp:564) */
uint32_t lo_step_seed(uint32_t moduleSeed, uint32_t frameIndex, uint32_t substep) {
    return lo_pcg_hash(moduleSeed ^ (frameIndex * 197u + substep));
}

/* SplitMix64 (Random.h:11-31): the constructor adds the golden gamma once,
 * next() adds it again before mixing. */
void lo_splitmix_init(uint64_t *state, uint64_t seed) { *state = seed + 0x9E3779B97F4A7C15ull; }

uint64_t lo_splitmix_next(uint64_t *state) {
    uint64_t z = (*state += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

void lo_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int lo_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------ */
/* kernel builders                                                     */
/* ------------------------------------------------------------------ */
static double core_value(int core, double x, double shellMu, double sigma) {
    if (core == LO_CORE_EXP) {
        if (x <= 0.0 || x >= 1.0) return 0.0;
        return exp(4.0 - 1.0 / (x * (1.0 - x)));
    }
    double t = (x - shellMu) / sigma;
    return exp(-0.5 * t * t);
}

int lo_kernel_half(const lo_kernel *k) { return (int)ceil(k->radius - 1e-9); }

/* unnormalised kernel value at integer offset (dx, dy) */
double lo_kernel_raw(const lo_kernel *k, int dx, int dy) {
    double rr = sqrt((double)dx * dx + (double)dy * dy);
    if (k->shape == LO_SHAPE_SHELL) {
        /* buildKernelTexture (Lenia.cpp:421-428): r normalised by R, 0<r<=1 */
        double r = rr / k->radius;
        if (!(r > 0.0 && r <= 1.0)) return 0.0;
        return core_value(k->core, r, k->shellMu, k->shellSigma);
    }
    /* buildRingKernelImage (Lenia.cpp:109-116): B shells, centre 0.5 */
    if (!(rr > 0.0 && rr <= k->radius) || k->nBetas <= 0) return 0.0;
    int B = k->nBetas;
    double u = (rr / k->radius) * (double)B;
    int i = (int)u;
    if (i > B - 1) i = B - 1;
    double frac = u - (double)i;
    if (k->core == LO_CORE_EXP) return k->betas[i] * core_value(LO_CORE_EXP, frac, 0.5, 1.0);
    double t = (frac - 0.5) / (k->shellSigma * (double)B);
    return k->betas[i] * exp(-0.5 * t * t);
}

int lo_kernel_window(const lo_kernel *k, int half, double *out) {
    int n = 2 * half + 1;
    double sum = 0.0;
    for (int y = 0; y < n; ++y)
        for (int x = 0; x < n; ++x) {
            double w = lo_kernel_raw(k, x - half, y - half);
            out[y * n + x] = w;
            sum += w;
        }
    if (sum > 0.0)
        for (int i = 0; i < n * n; ++i) out[i] /= sum;
    return sum > 0.0 ? 0 : -1;
}

/* wrap-centred placement (Lenia.cpp:103-108): dx > W/2 -> dx - W */
void lo_kernel_image(const lo_kernel *k, int W, int H, double *out) {
    double sum = 0.0;
    for (int y = 0; y < H; ++y) {
        int dy = y > H / 2 ? y - H : y;
        for (int x = 0; x < W; ++x) {
            int dx = x > W / 2 ? x - W : x;
            double w = lo_kernel_raw(k, dx, dy);
            out[(size_t)y * W + x] = w;
            sum += w;
        }
    }
    if (sum > 0.0)
        for (size_t i = 0; i < (size_t)W * H; ++i) out[i] /= sum;
}

/* ------------------------------------------------------------------ */
/* init                                                                */
/* ------------------------------------------------------------------ */
/* seeded uniform[0,1) noise inside a rectangle, zero elsewhere.  The
 * values are rand01(x, y, TAG_PATCH, seed) at global cell coordinates, the
 * same per-pixel hash the reference's shaders use (Random.metal:16-19). */
void lo_init_patch(int W, int H, uint32_t seed, int x0, int y0, int pw, int ph, double *out) {
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            int in = x >= x0 && x < x0 + pw && y >= y0 && y < y0 + ph;
            out[(size_t)y * W + x] =
                in ? lo_rand01((uint32_t)x, (uint32_t)y, LO_TAG_PATCH, seed) : 0.0;
        }
}

/* leniaValueNoise (Lenia.metal:32-41) */
static double value_noise(double px, double py, uint32_t tag, uint32_t seed) {
    double ix = floor(px) + 512.0, iy = floor(py) + 512.0;
    double fx = px - floor(px), fy = py - floor(py);
    fx = fx * fx * (3.0 - 2.0 * fx);
    fy = fy * fy * (3.0 - 2.0 * fy);
    uint32_t ux = (uint32_t)ix, uy = (uint32_t)iy;
    double a = lo_rand01(ux, uy, tag, seed);
    double b = lo_rand01(ux + 1u, uy, tag, seed);
    double c = lo_rand01(ux, uy + 1u, tag, seed);
    double d = lo_rand01(ux + 1u, uy + 1u, tag, seed);
    double ab = a + (b - a) * fx, cd = c + (d - c) * fx;
    return ab + (cd - ab) * fy;
}

/* leniaInit (Lenia.metal:43-61) */
void lo_init_value_noise(int W, int H, uint32_t seed, double cov, double scale, double *out) {
    double s = scale > 2.0 ? scale : 2.0;
    double c = cov > 1e-3 ? cov : 1e-3;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double px = x / s, py = y / s;
            double v = 0.55 * value_noise(px, py, 11u, seed) +
                       0.30 * value_noise(px * 2.13, py * 2.13, 23u, seed) +
                       0.15 * value_noise(px * 4.31, py * 4.31, 37u, seed);
            v = (v - (1.0 - cov)) / c;
            v = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
            out[(size_t)y * W + x] = v * 0.9;
        }
}

/* ------------------------------------------------------------------ */
/* organism stamps: leniaStampInit (Lenia.metal:242-303).  Stamp i sits  */
/* at (rand01(i,0,907), rand01(0,i,911)) * (W,H); with rotation on, its  */
/* orientation is one of 8 (rotate 0/90/180/270 x flip-x) from          */
/* rand01(i,1,919).  A cell takes the largest organism value of the     */
/* stamps covering it (nearest organism cell, toroidal offsets).  The    */
/* geometry is FP32 as in the shader, so cell boundaries match it.      */
/* ------------------------------------------------------------------ */
static void stamp_inverse(float *dx, float *dy, uint32_t variant) {
    float x = *dx, y = *dy;
    if (variant & 4u) x = -x;
    uint32_t rot = variant & 3u;
    float nx = x, ny = y;
    if (rot == 1u) { nx = y; ny = -x; }
    else if (rot == 2u) { nx = -x; ny = -y; }
    else if (rot == 3u) { nx = -y; ny = x; }
    *dx = nx;
    *dy = ny;
}

void lo_init_stamp(int W, int H, const double *org, int ow, int oh, uint32_t count, uint32_t rotate,
                   uint32_t seed, double *out) {
    uint32_t n = count < 32u ? count : 32u;
    float cx0 = (float)ow * 0.5f, cy0 = (float)oh * 0.5f;
    for (int y = 0; y < H; ++y)
        for (int x = 0; x < W; ++x) {
            double v = 0.0;
            for (uint32_t i = 0; i < n; ++i) {
                float cx = (float)lo_rand01(i, 0u, 907u, seed) * (float)W;
                float cy = (float)lo_rand01(0u, i, 911u, seed) * (float)H;
                uint32_t variant = 0u;
                if (rotate) {
                    uint32_t r = (uint32_t)((float)lo_rand01(i, 1u, 919u, seed) * 8.0f);
                    variant = r < 7u ? r : 7u;
                }
                float dx = (float)x + 0.5f - cx, dy = (float)y + 0.5f - cy;
                dx -= (float)W * roundf(dx / (float)W);
                dy -= (float)H * roundf(dy / (float)H);
                stamp_inverse(&dx, &dy, variant);
                int lx = (int)floorf(dx + cx0), ly = (int)floorf(dy + cy0);
                if (lx >= 0 && lx < ow && ly >= 0 && ly < oh) {
                    double o = org[(size_t)ly * ow + lx];
                    if (o > v) v = o;
                }
            }
            out[(size_t)y * W + x] = v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
        }
}

/* ------------------------------------------------------------------ */
/* colour layer: colorMapField / colorMapFieldScaled                   */
/* (ColorMap.metal:17-75): t = clamp(v*scale + bias, 0, 1);            */
/* rgb = max(a + b cos(2 pi (c t + d)), 0) * exposure; alpha = t.  An   */
/* R32F texel reads (r, 0, 0, 1).  Other sizes: linear filter at the    */
/* output pixel centre, repeat addressing.                              */
/* ------------------------------------------------------------------ */
void lo_colormap(const double *field, int srcW, int srcH, const double *a, const double *b,
                 const double *c, const double *d, double inputScale, double inputBias,
                 double exposure, int channel, double *rgba, int dstW, int dstH) {
    const double twoPi = 6.283185307179586;
    for (int y = 0; y < dstH; ++y)
        for (int x = 0; x < dstW; ++x) {
            double v;
            if (srcW == dstW && srcH == dstH) {
                v = field[(size_t)y * srcW + x];
            } else {
                double sx = (x + 0.5) / dstW * srcW - 0.5, sy = (y + 0.5) / dstH * srcH - 0.5;
                double fx0 = floor(sx), fy0 = floor(sy);
                double ax = sx - fx0, ay = sy - fy0;
                int x0 = (int)fx0 % srcW, y0 = (int)fy0 % srcH;
                if (x0 < 0) x0 += srcW;
                if (y0 < 0) y0 += srcH;
                int x1 = (x0 + 1) % srcW, y1 = (y0 + 1) % srcH;
                double v00 = field[(size_t)y0 * srcW + x0], v01 = field[(size_t)y0 * srcW + x1];
                double v10 = field[(size_t)y1 * srcW + x0], v11 = field[(size_t)y1 * srcW + x1];
                double top = v00 + ax * (v01 - v00), bot = v10 + ax * (v11 - v10);
                v = top + ay * (bot - top);
            }
            double sel = channel == 0 ? v : (channel >= 3 ? 1.0 : 0.0);
            double t = sel * inputScale + inputBias;
            t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
            double *o = rgba + 4 * ((size_t)y * dstW + x);
            for (int k = 0; k < 3; ++k) {
                double e = a[k] + b[k] * cos(twoPi * (c[k] * t + d[k]));
                o[k] = (e > 0.0 ? e : 0.0) * exposure;
            }
            o[3] = t;
        }
}

/* ------------------------------------------------------------------ */
/* direct potential: leniaStep's gather (Lenia.metal:73-83)             */
/* u(x,y) = sum_{dy,dx} K[dy][dx] * A[wrap(y+dy)][wrap(x+dx)]           */
/* ------------------------------------------------------------------ */
static inline int wrapi(int p, int n) { /* CommonTypes.metal:39-45 */
    int r = p % n;
    return r < 0 ? r + n : r;
}

void lo_potential_direct(int W, int H, const double *A, const double *win, int half, double *u) {
    int n = 2 * half + 1;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            double acc = 0.0;
            for (int dy = -half; dy <= half; ++dy) {
                const double *krow = win + (size_t)(dy + half) * n;
                const double *arow = A + (size_t)wrapi(y + dy, H) * W;
                for (int dx = -half; dx <= half; ++dx) {
                    double w = krow[dx + half];
                    if (w == 0.0) continue;
                    acc += w * arow[wrapi(x + dx, W)];
                }
            }
            u[(size_t)y * W + x] = acc;
        }
    }
}

/* ------------------------------------------------------------------ */
/* FP64 mixed-radix FFT (any N; smallest-prime-factor recursion).       */
/* forward X[k] = sum x[n] e^{-2 pi i k n / N}, inverse uses e^{+...},   */
/* both unscaled (FFT.metal:18-34 convention before its 1/N stage).      */
/* ------------------------------------------------------------------ */
typedef double complex cplx;

static int smallest_factor(int n) {
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    for (int p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

static void fft_rec(const cplx *in, int istride, cplx *out, int N, int s, const cplx *T, int inv,
                    cplx *scratch) {
    if (N == 1) {
        out[0] = in[0];
        return;
    }
    int p = smallest_factor(N), m = N / p;
    for (int r = 0; r < p; ++r)
        fft_rec(in + (size_t)r * istride, istride * p, out + (size_t)r * m, m, s * p, T, inv,
                scratch + N);
    memcpy(scratch, out, sizeof(cplx) * N);
    for (int k = 0; k < N; ++k) {
        cplx acc = 0;
        for (int r = 0; r < p; ++r) {
            int e = (int)(((long long)r * k) % N);
            cplx w = T[(size_t)e * s];
            if (inv) w = conj(w);
            acc += w * scratch[r * m + (k % m)];
        }
        out[k] = acc;
    }
}

static cplx *twiddle_table(int N) {
    cplx *T = (cplx *)malloc(sizeof(cplx) * N);
    for (int j = 0; j < N; ++j) {
        double a = -2.0 * M_PI * (double)j / (double)N;
        T[j] = cos(a) + I * sin(a);
    }
    return T;
}

static void fft_line(cplx *x, int N, const cplx *T, int inv, cplx *tmp, cplx *scratch) {
    memcpy(tmp, x, sizeof(cplx) * N);
    fft_rec(tmp, 1, x, N, 1, T, inv, scratch);
}

void lo_fft(int N, double *re, double *im, int inverse) {
    cplx *T = twiddle_table(N);
    cplx *x = (cplx *)malloc(sizeof(cplx) * N * 4);
    for (int i = 0; i < N; ++i) x[i] = re[i] + I * im[i];
    fft_line(x, N, T, inverse, x + N, x + 2 * N);
    for (int i = 0; i < N; ++i) {
        re[i] = creal(x[i]);
        im[i] = cimag(x[i]);
    }
    free(x);
    free(T);
}

static void fft2_c(int W, int H, cplx *d, int inv) {
    cplx *TW = twiddle_table(W), *TH = twiddle_table(H);
#pragma omp parallel
    {
        cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)(W > H ? W : H) * 4);
        int L = W > H ? W : H;
#pragma omp for schedule(static)
        for (int y = 0; y < H; ++y) fft_line(d + (size_t)y * W, W, TW, inv, buf, buf + L);
#pragma omp for schedule(static)
        for (int x = 0; x < W; ++x) {
            cplx *col = buf + 3 * L;
            for (int y = 0; y < H; ++y) col[y] = d[(size_t)y * W + x];
            fft_line(col, H, TH, inv, buf, buf + L);
            for (int y = 0; y < H; ++y) d[(size_t)y * W + x] = col[y];
        }
        free(buf);
    }
    free(TW);
    free(TH);
}

void lo_fft2(int W, int H, double *re, double *im, int inverse) {
    size_t n = (size_t)W * H;
    cplx *d = (cplx *)malloc(sizeof(cplx) * n);
    for (size_t i = 0; i < n; ++i) d[i] = re[i] + I * im[i];
    fft2_c(W, H, d, inverse);
    for (size_t i = 0; i < n; ++i) {
        re[i] = creal(d[i]);
        im[i] = cimag(d[i]);
    }
    free(d);
}

/* potential by the convolution theorem with a precomputed kernel spectrum:
 * u = Re(IFFT(FFT(A) * FFT(img))) / (W*H).  Circular convolution with the
 * wrap-centred image equals the direct gather because every kernel is even
 * (Lenia.cpp:576-616 computes the same quantity on the GPU). */
static void potential_fft_spec(int W, int H, const double *A, const cplx *kspec, double *u) {
    size_t n = (size_t)W * H;
    cplx *d = (cplx *)malloc(sizeof(cplx) * n);
    for (size_t i = 0; i < n; ++i) d[i] = A[i];
    fft2_c(W, H, d, 0);
    for (size_t i = 0; i < n; ++i) d[i] *= kspec[i];
    fft2_c(W, H, d, 1);
    double s = 1.0 / (double)n;
    for (size_t i = 0; i < n; ++i) u[i] = creal(d[i]) * s;
    free(d);
}

static cplx *kernel_spectrum(const lo_kernel *k, int W, int H) {
    size_t n = (size_t)W * H;
    double *img = (double *)malloc(sizeof(double) * n);
    lo_kernel_image(k, W, H, img);
    cplx *d = (cplx *)malloc(sizeof(cplx) * n);
    for (size_t i = 0; i < n; ++i) d[i] = img[i];
    fft2_c(W, H, d, 0);
    free(img);
    return d;
}

void lo_potential_fft(int W, int H, const double *A, const double *img, double *u) {
    size_t n = (size_t)W * H;
    cplx *ks = (cplx *)malloc(sizeof(cplx) * n);
    for (size_t i = 0; i < n; ++i) ks[i] = img[i];
    fft2_c(W, H, ks, 0);
    potential_fft_spec(W, H, A, ks, u);
    free(ks);
}

/* ------------------------------------------------------------------ */
/* growth: leniaStep tail (Lenia.metal:85-127) and leniaGrowthMulti     */
/* (Lenia.metal:161-232).  For one kernel of weight 1 both formulas      */
/* coincide, so one routine serves both passes.                         */
/* ------------------------------------------------------------------ */
static double metal_round(double v) { return v < 0.0 ? -floor(-v + 0.5) : floor(v + 0.5); }

static void growth_impl(const lo_world *w, const double *A, const double *const *u,
                        uint32_t seed, uint32_t frameIndex, double *out, int clip) {
    const int W = w->width, H = w->height, K = w->nKernels;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            size_t i = (size_t)y * W + x;
            double a = A[i];
            double muJ = 0.0;
            if (w->muJitter > 0.0) {
                double j = lo_rand01((uint32_t)x, (uint32_t)y, 401u + frameIndex % 613u, seed) - 0.5;
                muJ = j * w->muJitter * 0.05;
            }
            double sumWb = 0.0, sumW = 0.0, sumStd = 0.0;
            for (int k = 0; k < K; ++k) {
                double t = (u[k][i] - (w->k[k].mu + muJ)) / w->k[k].sigma;
                double bell = exp(-0.5 * t * t);
                sumWb += w->k[k].weight * bell;
                sumW += w->k[k].weight;
                sumStd += w->k[k].weight * (2.0 * bell - 1.0);
            }
            double a2;
            if (w->growthMode == 1) {
                double target = sumW > 1e-6 ? sumWb / sumW : a;
                a2 = a + w->dt * (target - a);
            } else {
                a2 = a + w->dt * sumStd;
            }
            if (w->noiseAmount > 0.0) {
                double nz = lo_rand01((uint32_t)x, (uint32_t)y, 503u, seed) - 0.5;
                a2 += nz * w->noiseAmount * 0.03;
            }
            if (w->injectAmount > 0.0) {
                double uvx = (x + 0.5) / W, uvy = (y + 0.5) / H;
                double aspect = (double)W / (double)H;
                uint32_t cycle = frameIndex / 12u;
                double blobR = (double)w->radius * 0.9 / (double)H;
                for (int b = 0; b < w->injectCount; ++b) {
                    double cx = lo_rand01((uint32_t)b, cycle, 811u, seed);
                    double cy = lo_rand01(cycle, (uint32_t)b, 823u, seed);
                    double dx = uvx - cx, dy = uvy - cy;
                    dx -= metal_round(dx);
                    dy -= metal_round(dy);
                    dx *= aspect;
                    double r2 = (dx * dx + dy * dy) / (blobR * blobR);
                    double bump = exp(-r2 * 0.5);
                    double t = bump * w->injectAmount;
                    t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
                    a2 = a2 + (0.6 - a2) * t;
                }
            }
            out[i] = clip ? (a2 < 0.0 ? 0.0 : (a2 > 1.0 ? 1.0 : a2)) : a2;
        }
    }
}

void lo_growth(const lo_world *w, const double *A, const double *const *u, uint32_t seed,
               uint32_t frameIndex, double *out) {
    growth_impl(w, A, u, seed, frameIndex, out, 1);
}

/* ------------------------------------------------------------------ */
/* one step                                                            */
/* ------------------------------------------------------------------ */
typedef struct {
    int half[4];
    double *win[4];
    cplx *spec[4];
} prepared;

static void prepare(const lo_world *w, prepared *p, int want_fft) {
    memset(p, 0, sizeof(*p));
    for (int k = 0; k < w->nKernels; ++k) {
        p->half[k] = lo_kernel_half(&w->k[k]);
        int n = 2 * p->half[k] + 1;
        p->win[k] = (double *)malloc(sizeof(double) * n * n);
        lo_kernel_window(&w->k[k], p->half[k], p->win[k]);
        if (want_fft) p->spec[k] = kernel_spectrum(&w->k[k], w->width, w->height);
    }
}

static void unprepare(const lo_world *w, prepared *p) {
    for (int k = 0; k < w->nKernels; ++k) {
        free(p->win[k]);
        free(p->spec[k]);
    }
}

static void potentials(const lo_world *w, const prepared *p, const double *A, double **u,
                       int use_fft) {
    for (int k = 0; k < w->nKernels; ++k) {
        if (use_fft)
            potential_fft_spec(w->width, w->height, A, p->spec[k], u[k]);
        else
            lo_potential_direct(w->width, w->height, A, p->win[k], p->half[k], u[k]);
    }
}

int lo_step(const lo_world *w, const double *A, double *out, int use_fft, uint32_t seed,
            uint32_t frameIndex) {
    if (w->nKernels < 1 || w->nKernels > 4) return -1;
    size_t n = (size_t)w->width * w->height;
    prepared p;
    prepare(w, &p, use_fft);
    double *u[4];
    for (int k = 0; k < w->nKernels; ++k) u[k] = (double *)malloc(sizeof(double) * n);
    potentials(w, &p, A, u, use_fft);
    lo_growth(w, A, (const double *const *)u, seed, frameIndex, out);
    for (int k = 0; k < w->nKernels; ++k) free(u[k]);
    unprepare(w, &p);
    return 0;
}

/* T steps with zero noise/jitter/inject (seed-independent) */
int lo_rollout(const lo_world *w, const double *A0, int T, int use_fft, double *AT) {
    if (w->nKernels < 1 || w->nKernels > 4) return -1;
    size_t n = (size_t)w->width * w->height;
    prepared p;
    prepare(w, &p, use_fft);
    double *u[4];
    for (int k = 0; k < w->nKernels; ++k) u[k] = (double *)malloc(sizeof(double) * n);
    double *cur = (double *)malloc(sizeof(double) * n);
    memcpy(cur, A0, sizeof(double) * n);
    for (int t = 0; t < T; ++t) {
        potentials(w, &p, cur, u, use_fft);
        lo_growth(w, cur, (const double *const *)u, 0u, 0u, AT);
        memcpy(cur, AT, sizeof(double) * n);
    }
    memcpy(AT, cur, sizeof(double) * n);
    free(cur);
    for (int k = 0; k < w->nKernels; ++k) free(u[k]);
    unprepare(w, &p);
    return 0;
}

/* T-step rollout on one smooth piece of the clipped map (test helper, */
/* no reference counterpart).  record != 0: the plain rollout, storing  */
/* each step's clip state in mask[t*n + i] (0 inside, 1 clipped to 0,   */
/* 2 clipped to 1).  record == 0: the same steps with those states      */
/* imposed (0 -> the unclipped pre-clip value, 1 -> 0, 2 -> 1), i.e. the */
/* smooth branch the recorded trajectory lies on; finite differences of */
/* it are free of clip kinks, and equal the gradient the adjoint takes. */
int lo_rollout_piece(const lo_world *w, const double *A0, int T, unsigned char *mask, int record,
                     double *AT) {
    if (w->nKernels < 1 || w->nKernels > 4) return -1;
    size_t n = (size_t)w->width * w->height;
    prepared p;
    prepare(w, &p, 1);
    double *u[4];
    for (int k = 0; k < w->nKernels; ++k) u[k] = (double *)malloc(sizeof(double) * n);
    double *cur = (double *)malloc(sizeof(double) * n);
    memcpy(cur, A0, sizeof(double) * n);
    for (int t = 0; t < T; ++t) {
        unsigned char *m = mask + n * (size_t)t;
        potentials(w, &p, cur, u, 1);
        growth_impl(w, cur, (const double *const *)u, 0u, 0u, AT, 0);
        for (size_t i = 0; i < n; ++i) {
            if (record) {
                m[i] = AT[i] <= 0.0 ? 1 : (AT[i] >= 1.0 ? 2 : 0);
                cur[i] = AT[i] < 0.0 ? 0.0 : (AT[i] > 1.0 ? 1.0 : AT[i]);
            } else {
                cur[i] = m[i] == 1 ? 0.0 : (m[i] == 2 ? 1.0 : AT[i]);
            }
        }
    }
    memcpy(AT, cur, sizeof(double) * n);
    free(cur);
    for (int k = 0; k < w->nKernels; ++k) free(u[k]);
    unprepare(w, &p);
    return 0;
}

/* ------------------------------------------------------------------ */
/* FP64 reference gradient of L = sum(gT * A_T) for a T-step rollout in  */
/* growth mode 0 without noise, jitter or inject.                        */
/*   pre = A + dt * sum_k w_k (2 b_k - 1),  A' = clip(pre, 0, 1)         */
/*   b_k = exp(-(u_k - mu_k)^2 / (2 sigma_k^2)),  u_k = K_k (x) A         */
/* The clip derivative is 1 strictly inside (0,1) and 0 elsewhere.       */
/* The kernel is K = raw / S with raw(d) = beta_{i(d)} c(d); its gradient */
/* reaches the betas through the normalisation S = sum_j beta_j C_j.      */
/* ------------------------------------------------------------------ */
int lo_rollout_grad(const lo_world *w, const double *A0, int T, const double *gT, double *gA0,
                    double *gMu, double *gSigma, double *gWeight, double *gBetas, double *gDt) {
    if (w->nKernels < 1 || w->nKernels > 4 || w->growthMode != 0) return -1;
    const int W = w->width, H = w->height, K = w->nKernels;
    size_t n = (size_t)W * H;
    prepared p;
    prepare(w, &p, 1);
    double *states = (double *)malloc(sizeof(double) * n * (size_t)(T + 1));
    memcpy(states, A0, sizeof(double) * n);
    double *u[4], *gu[4];
    for (int k = 0; k < K; ++k) {
        u[k] = (double *)malloc(sizeof(double) * n);
        gu[k] = (double *)malloc(sizeof(double) * n);
    }
    for (int t = 0; t < T; ++t) {
        potentials(w, &p, states + n * t, u, 1);
        lo_growth(w, states + n * t, (const double *const *)u, 0u, 0u, states + n * (t + 1));
    }
    double *g = (double *)malloc(sizeof(double) * n);
    double *gn = (double *)malloc(sizeof(double) * n);
    double *tmp = (double *)malloc(sizeof(double) * n);
    memcpy(g, gT, sizeof(double) * n);
    double gK[4][4]; /* per kernel, per shell: sum_d c(d) gKimg(d) */
    double gKK[4];   /* per kernel: sum_d K(d) gKimg(d) */
    memset(gK, 0, sizeof(gK));
    memset(gKK, 0, sizeof(gKK));
    for (int k = 0; k < K; ++k) {
        gMu[k] = gSigma[k] = gWeight[k] = 0.0;
        for (int j = 0; j < 4; ++j) gBetas[4 * k + j] = 0.0;
    }
    *gDt = 0.0;
    cplx *fa = (cplx *)malloc(sizeof(cplx) * n);
    cplx *fg = (cplx *)malloc(sizeof(cplx) * n);
    /* per-kernel shell images c_j (unnormalised raw with beta_j -> 1) */
    double *shellImg[4][4];
    double shellSum[4][4];
    double rawSum[4];
    for (int k = 0; k < K; ++k) {
        lo_kernel kk = w->k[k];
        int B = kk.shape == LO_SHAPE_RING ? kk.nBetas : 1;
        rawSum[k] = 0.0;
        for (int j = 0; j < 4; ++j) shellImg[k][j] = NULL;
        for (int j = 0; j < B; ++j) {
            lo_kernel one = kk;
            for (int q = 0; q < 4; ++q) one.betas[q] = (q == j) ? 1.0 : 0.0;
            double *img = (double *)malloc(sizeof(double) * n);
            shellSum[k][j] = 0.0;
            for (int y = 0; y < H; ++y) {
                int dy = y > H / 2 ? y - H : y;
                for (int x = 0; x < W; ++x) {
                    int dx = x > W / 2 ? x - W : x;
                    double v = lo_kernel_raw(&one, dx, dy);
                    img[(size_t)y * W + x] = v;
                    shellSum[k][j] += v;
                }
            }
            shellImg[k][j] = img;
            rawSum[k] += (kk.shape == LO_SHAPE_RING ? kk.betas[j] : 1.0) * shellSum[k][j];
        }
    }
    double *kimg = (double *)malloc(sizeof(double) * n);
    for (int t = T - 1; t >= 0; --t) {
        const double *A = states + n * t;
        potentials(w, &p, A, u, 1);
        double gdt = 0.0;
        double gmu[4] = {0}, gsg[4] = {0}, gw[4] = {0};
        for (size_t i = 0; i < n; ++i) {
            double sumStd = 0.0, b[4];
            for (int k = 0; k < K; ++k) {
                double t1 = (u[k][i] - w->k[k].mu) / w->k[k].sigma;
                b[k] = exp(-0.5 * t1 * t1);
                sumStd += w->k[k].weight * (2.0 * b[k] - 1.0);
            }
            double pre = A[i] + w->dt * sumStd;
            double gp = (pre > 0.0 && pre < 1.0) ? g[i] : 0.0;
            gn[i] = gp;
            gdt += gp * sumStd;
            for (int k = 0; k < K; ++k) {
                double mu = w->k[k].mu, sg = w->k[k].sigma, wk = w->k[k].weight;
                double d = u[k][i] - mu;
                double c = gp * w->dt * wk * 2.0 * b[k];
                gmu[k] += c * d / (sg * sg);
                gsg[k] += c * d * d / (sg * sg * sg);
                gw[k] += gp * w->dt * (2.0 * b[k] - 1.0);
                gu[k][i] = -c * d / (sg * sg);
            }
        }
        *gDt += gdt;
        for (int k = 0; k < K; ++k) {
            gMu[k] += gmu[k];
            gSigma[k] += gsg[k];
            gWeight[k] += gw[k];
            /* dL/dA += K (x) gu  (K even: the adjoint of the gather is the
             * same gather) */
            potential_fft_spec(W, H, gu[k], p.spec[k], tmp);
            for (size_t i = 0; i < n; ++i) gn[i] += tmp[i];
            /* kernel image gradient gKimg(d) = sum_x gu(x) A(x+d):
             * correlation, via FFT: IFFT(conj(FFT(gu)) * FFT(A)) */
            for (size_t i = 0; i < n; ++i) {
                fa[i] = A[i];
                fg[i] = gu[k][i];
            }
            fft2_c(W, H, fa, 0);
            fft2_c(W, H, fg, 0);
            for (size_t i = 0; i < n; ++i) fa[i] = conj(fg[i]) * fa[i];
            fft2_c(W, H, fa, 1);
            double s = 1.0 / (double)n;
            lo_kernel_image(&w->k[k], W, H, kimg);
            double accKK = 0.0;
            for (size_t i = 0; i < n; ++i) accKK += kimg[i] * creal(fa[i]) * s;
            gKK[k] += accKK;
            int B = w->k[k].shape == LO_SHAPE_RING ? w->k[k].nBetas : 1;
            for (int j = 0; j < B; ++j) {
                double acc = 0.0;
                for (size_t i = 0; i < n; ++i) acc += shellImg[k][j][i] * creal(fa[i]) * s;
                gK[k][j] += acc;
            }
        }
        memcpy(g, gn, sizeof(double) * n);
    }
    memcpy(gA0, g, sizeof(double) * n);
    for (int k = 0; k < K; ++k) {
        int B = w->k[k].shape == LO_SHAPE_RING ? w->k[k].nBetas : 1;
        for (int j = 0; j < B; ++j)
            gBetas[4 * k + j] = (gK[k][j] - shellSum[k][j] * gKK[k]) / rawSum[k];
        for (int j = 0; j < B; ++j) free(shellImg[k][j]);
    }
    free(kimg);
    free(fa);
    free(fg);
    free(g);
    free(gn);
    free(tmp);
    for (int k = 0; k < K; ++k) {
        free(u[k]);
        free(gu[k]);
    }
    free(states);
    unprepare(w, &p);
    return 0;
}
