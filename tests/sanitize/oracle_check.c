/* Drives the CPU oracle (oracle/oracle.c) under AddressSanitizer and
 * UndefinedBehaviorSanitizer: FFTs of mixed-radix lengths, the four PSF
 * types, RL planes at both bit depths with a warm start, odd sizes and
 * non-square kernels, the transform (including a z-subset) and the mask.
 * Prints a checksum. */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "oracle.h"

static double frand(unsigned *s)
{
    *s = *s * 1103515245u + 12345u;
    return (double)((*s >> 8) & 0xffff) / 65536.0;
}

int main(void)
{
    unsigned s = 1;
    double sum = 0.0;
    int ns[] = {1, 2, 12, 30, 97, 160, 2560};
    for (int i = 0; i < 7; ++i) {
        int n = ns[i];
        double *x = malloc(sizeof(double) * 2 * n);
        for (int k = 0; k < 2 * n; ++k) x[k] = frand(&s);
        if (orc_fft_c2c_f64(x, n, 0) == 0) sum += x[0];
        free(x);
    }
    for (int psf = 0; psf < 4; ++psf) {
        const int Z = 2, V = 3, dh = 21, dw = 17, kh = 7, kw = 5, bits = psf == 2 ? 8 : 16;
        uint16_t *data = malloc(sizeof(uint16_t) * Z * V * dh * dw);
        float *w = malloc(sizeof(float) * V * dh * dw), *k = malloc(sizeof(float) * V * kh * kw);
        double *init = malloc(sizeof(double) * Z * dh * dw), *est = malloc(sizeof(double) * Z * dh * dw);
        uint16_t *out = malloc(sizeof(uint16_t) * Z * dh * dw);
        for (int i = 0; i < Z * V * dh * dw; ++i) data[i] = (uint16_t)(frand(&s) * (bits == 8 ? 255 : 4000));
        for (int i = 0; i < V * dh * dw; ++i) w[i] = frand(&s) < 0.2 ? 0.0f : (float)frand(&s);
        for (int i = 0; i < V * kh * kw; ++i) k[i] = (float)frand(&s) + 0.01f;
        for (int i = 0; i < Z * dh * dw; ++i) init[i] = 50.0 + 100.0 * frand(&s);
        if (orc_rl_planes_f64(dw, dh, Z, V, bits, data, w, k, kw, kh, psf, 3, psf & 1 ? init : NULL, est, out, 2) == 0)
            for (int i = 0; i < Z * dh * dw; ++i) sum += est[i] + out[i];
        if (orc_rl_planes_f32(dw, dh, Z, V, bits, data, w, k, kw, kh, psf, 2, NULL, est, out, 1) == 0)
            for (int i = 0; i < Z * dh * dw; ++i) sum += est[i];
        free(data); free(w); free(k); free(init); free(est); free(out);
    }
    {
        const int w = 23, h = 19, d = 11, tw = 21, th = 17, td = 9;
        uint16_t *src = malloc(sizeof(uint16_t) * w * h * d), *dst = malloc(sizeof(uint16_t) * tw * th * td);
        float m[12] = {0.97f, 0.1f, 0.02f, 0.6f, -0.1f, 0.98f, 0.03f, -0.4f, 0.01f, -0.02f, 1.02f, 0.3f};
        float *mask = malloc(sizeof(float) * tw * th);
        for (int i = 0; i < w * h * d; ++i) src[i] = (uint16_t)(frand(&s) * 65535);
        orc_transform(src, w, h, d, 16, m, tw, th, 0, td, dst);
        for (int i = 0; i < tw * th * td; ++i) sum += dst[i];
        orc_transform(src, w, h, d, 8, m, tw, th, 3, 2, dst);
        for (int i = 0; i < tw * th * 2; ++i) sum += dst[i];
        orc_mask(w, h, d, tw, th, td, m, 4, 1.5f, mask);
        for (int i = 0; i < tw * th; ++i) sum += mask[i];
        free(src); free(dst); free(mask);
    }
    printf("oracle_check ok %.6e\n", sum);
    return isfinite(sum) ? 0 : 1;
}
