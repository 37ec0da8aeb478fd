/*
 * CPU FP64 oracle of the learned pilot network -- TEST INFRASTRUCTURE ONLY.
 * Only tests/ and __graft_entry__.smoke() may load it, as the reference result.
 *
 * Network (the node's design; no reference counterpart): input channels
 * [Y, U, V, 10*sigma] of the noisy opponent planes; L1 conv3x3 4->32 + ReLU;
 * L2..L5 conv3x3 32->32 + ReLU; L6 conv3x3 32->3; output = input[0..2] + L6.
 * Zero padding, stride 1. Parameters use the engine's flat layout: per layer
 * W[tap 0..8][CIN_PAD][COUT_PAD] then b[COUT_PAD] (tap = ky*3+kx, padded
 * entries ignored): CIN_PAD = {16,32,32,32,32,32}, COUT_PAD = {32,32,32,32,32,16}.
 * Loss = 0.5 * mean over (images, pixels, 3 channels) of (out - target)^2.
 * Pinned by central finite differences (tests/test_pilot_oracle.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NL 6
static const int CINP[NL] = {16, 32, 32, 32, 32, 32};
static const int COUTP[NL] = {32, 32, 32, 32, 32, 16};
static const int CINR[NL] = {4, 32, 32, 32, 32, 32};
static const int COUTR[NL] = {32, 32, 32, 32, 32, 3};

static size_t layer_off(int l) {
    size_t o = 0;
    for (int k = 0; k < l; ++k) o += (size_t)9 * CINP[k] * COUTP[k] + COUTP[k];
    return o;
}

size_t oracle_pilot_param_count(void) { return layer_off(NL); }

/* x: [ci][h][w] (real channels only) -> y: [co][h][w] pre-activation */
static void conv(const float* p, int l, const double* x, int w, int h, double* y) {
    const float* W = p + layer_off(l);
    const float* b = W + (size_t)9 * CINP[l] * COUTP[l];
    const size_t plane = (size_t)w * h;
    for (int co = 0; co < COUTR[l]; ++co)
        for (size_t i = 0; i < plane; ++i) y[co * plane + i] = b[co];
    for (int t = 0; t < 9; ++t) {
        const int dy = t / 3 - 1, dx = t % 3 - 1;
        for (int ci = 0; ci < CINR[l]; ++ci)
            for (int co = 0; co < COUTR[l]; ++co) {
                const double wv = W[((size_t)t * CINP[l] + ci) * COUTP[l] + co];
                for (int yy = 0; yy < h; ++yy) {
                    const int sy = yy + dy;
                    if (sy < 0 || sy >= h) continue;
                    for (int xx = 0; xx < w; ++xx) {
                        const int sx = xx + dx;
                        if (sx < 0 || sx >= w) continue;
                        y[co * plane + (size_t)yy * w + xx] += wv * x[ci * plane + (size_t)sy * w + sx];
                    }
                }
            }
    }
}

/* forward of one image; acts[l] (l = 0..5) receive the inputs of each layer
 * (post-ReLU), each sized 32 * plane; out: [3][plane] */
static void forward1(const float* p, const double* in, double sigma, int w, int h, double** acts,
                     double* out) {
    const size_t plane = (size_t)w * h;
    double* a0 = acts[0];
    for (int c = 0; c < 3; ++c) memcpy(a0 + c * plane, in + c * plane, plane * sizeof(double));
    for (size_t i = 0; i < plane; ++i) a0[3 * plane + i] = 10.0 * sigma;
    for (int l = 0; l < NL - 1; ++l) {
        conv(p, l, acts[l], w, h, acts[l + 1]);
        for (size_t i = 0; i < (size_t)COUTR[l] * plane; ++i)
            if (acts[l + 1][i] < 0.0) acts[l + 1][i] = 0.0;
    }
    conv(p, NL - 1, acts[NL - 1], w, h, out);
    for (size_t i = 0; i < 3 * plane; ++i) out[i] += in[i];
}

static double** alloc_acts(size_t plane) {
    double** a = (double**)malloc(sizeof(double*) * NL);
    for (int l = 0; l < NL; ++l) a[l] = (double*)calloc(32 * plane, sizeof(double));
    return a;
}

static void free_acts(double** a) {
    for (int l = 0; l < NL; ++l) free(a[l]);
    free(a);
}

void oracle_pilot_forward_f64(const float* params, const double* in, const double* sigma, int w,
                              int h, int n_img, double* out) {
    const size_t plane = (size_t)w * h;
    double** acts = alloc_acts(plane);
    for (int i = 0; i < n_img; ++i) forward1(params, in + (size_t)i * 3 * plane, sigma[i], w, h, acts,
                                             out + (size_t)i * 3 * plane);
    free_acts(acts);
}

double oracle_pilot_loss_f64(const float* params, const double* in, const double* target,
                             const double* sigma, int w, int h, int n_img) {
    const size_t plane = (size_t)w * h;
    double* out = (double*)malloc(sizeof(double) * 3 * plane * n_img);
    oracle_pilot_forward_f64(params, in, sigma, w, h, n_img, out);
    double s = 0.0;
    for (size_t i = 0; i < 3 * plane * n_img; ++i) {
        const double d = out[i] - target[i];
        s += d * d;
    }
    free(out);
    return 0.5 * s / (double)(3 * plane * n_img);
}

/* gradient of the loss w.r.t. every parameter (padded entries stay 0) */
double oracle_pilot_grad_f64(const float* params, const double* in, const double* target,
                             const double* sigma, int w, int h, int n_img, double* grad) {
    const size_t plane = (size_t)w * h;
    const size_t np = oracle_pilot_param_count();
    memset(grad, 0, np * sizeof(double));
    double** acts = alloc_acts(plane);
    double* out = (double*)malloc(sizeof(double) * 3 * plane);
    double* gy = (double*)malloc(sizeof(double) * 32 * plane);
    double* gx = (double*)malloc(sizeof(double) * 32 * plane);
    const double norm = 1.0 / (double)(3 * plane * n_img);
    double loss = 0.0;
    for (int img = 0; img < n_img; ++img) {
        const double* xin = in + (size_t)img * 3 * plane;
        const double* tg = target + (size_t)img * 3 * plane;
        forward1(params, xin, sigma[img], w, h, acts, out);
        memset(gy, 0, sizeof(double) * 32 * plane);
        for (size_t i = 0; i < 3 * plane; ++i) {
            const double d = out[i] - tg[i];
            loss += d * d;
            gy[i] = d * norm; /* dL/d(pre-activation of L6) */
        }
        for (int l = NL - 1; l >= 0; --l) {
            const float* W = params + layer_off(l);
            double* gW = grad + layer_off(l);
            double* gb = gW + (size_t)9 * CINP[l] * COUTP[l];
            const double* x = acts[l];
            for (int co = 0; co < COUTR[l]; ++co)
                for (size_t i = 0; i < plane; ++i) gb[co] += gy[co * plane + i];
            memset(gx, 0, sizeof(double) * 32 * plane);
            for (int t = 0; t < 9; ++t) {
                const int dy = t / 3 - 1, dx = t % 3 - 1;
                for (int ci = 0; ci < CINR[l]; ++ci)
                    for (int co = 0; co < COUTR[l]; ++co) {
                        const size_t wi = ((size_t)t * CINP[l] + ci) * COUTP[l] + co;
                        double acc = 0.0;
                        for (int yy = 0; yy < h; ++yy) {
                            const int sy = yy + dy;
                            if (sy < 0 || sy >= h) continue;
                            for (int xx = 0; xx < w; ++xx) {
                                const int sx = xx + dx;
                                if (sx < 0 || sx >= w) continue;
                                const double g = gy[co * plane + (size_t)yy * w + xx];
                                acc += g * x[ci * plane + (size_t)sy * w + sx];
                                gx[ci * plane + (size_t)sy * w + sx] += g * W[wi];
                            }
                        }
                        gW[wi] += acc;
                    }
            }
            if (l == 0) break;
            /* through the ReLU that produced acts[l] */
            for (int ci = 0; ci < CINR[l]; ++ci)
                for (size_t i = 0; i < plane; ++i)
                    gy[ci * plane + i] = x[ci * plane + i] > 0.0 ? gx[ci * plane + i] : 0.0;
        }
    }
    free(out);
    free(gy);
    free(gx);
    free_acts(acts);
    return 0.5 * loss * norm;
}
