/*
 * CPU oracle (test infrastructure only; never linked into the product).
 * See oracle.h for the contract and for what pins each function.
 */
#define _GNU_SOURCE
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* fmvd_deconvolve.cpp:22-26 (iAlignUp) */
int orc_ialign_up(int a, int b) { return (a % b != 0) ? (a - a % b + b) : a; }

/* fmvd_deconvolve.cpp:28-51 (snapTransformSize): align to 16; keep powers of
 * two; otherwise the next power of two if it is <= 1024, else align to 512. */
int orc_snap_transform_size(int n)
{
    int hibit;
    unsigned int lo, hi;
    n = orc_ialign_up(n, 16);
    for (hibit = 31; hibit >= 0; hibit--)
        if ((unsigned int)n & (1U << hibit)) break;
    lo = 1U << hibit;
    if (lo == (unsigned int)n) return n;
    hi = 1U << (hibit + 1);
    if (hi <= 1024) return (int)hi;
    return orc_ialign_up(n, 512);
}

int orc_pad_index(int i, int n, int half)
{
    if (i < n) return i;
    if (i < n + half) return n - 1;
    return 0;
}

/* ---- precision-generic FFT + RL core ---- */
#define REAL double
#define SFX f64
#include "oracle_real.inc"
#undef REAL
#undef SFX

#define REAL float
#define SFX f32
#include "oracle_real.inc"
#undef REAL
#undef SFX

int orc_fft2_c2c_f64(double *data, int h, int w, int inverse)
{
    plan_t_f64 ph, pw;
    if (plan_init_f64(&ph, h) || plan_init_f64(&pw, w)) return -1;
    fft2_run_f64(&ph, &pw, data, h, w, inverse, 1);
    plan_free_f64(&ph);
    plan_free_f64(&pw);
    return 0;
}

/* fmvd_deconvolve.cpp:53-68 */
void orc_normalize_f32(float *k, int len)
{
    double sum = 0;
    int i;
    for (i = 0; i < len; i++) sum += k[i];
    for (i = 0; i < len; i++) k[i] /= (float)sum;
}

static void normalize_f64(double *k, int len)
{
    double sum = 0;
    int i;
    for (i = 0; i < len; i++) sum += k[i];
    for (i = 0; i < len; i++) k[i] /= sum;
}

/* fmvd_deconvolve.cpp:70-79 */
void orc_flip_kernel_f64(const double *in, double *out, int kw, int kh)
{
    int x, y;
    for (y = 0; y < kh; y++)
        for (x = 0; x < kw; x++)
            out[y * kw + x] = in[(kh - y - 1) * kw + (kw - x - 1)];
}

/* fmvd_deconvolve.cpp:90-137.  The reference pads the data with the clamp
 * map (half-widths kh/2, kw/2 in their natural order, :118), pads the kernel
 * centred at (0,0) with wraparound (fmvd_deconvolve_cuda.cu:27-37), multiplies
 * the spectra and crops the top-left dh x dw.  That is the circular
 * convolution  res[y][x] = sum_ij K[i][j] Dp[(y+kY-i) mod H][(x+kX-j) mod W],
 * evaluated here directly in FP64. */
void orc_convolve_single_plane(double *data, int dw, int dh,
                               const double *kernel, int kw, int kh)
{
    const int ffth = orc_snap_transform_size(dh + kh - 1);
    const int fftw = orc_snap_transform_size(dw + kw - 1);
    const int ky = kh / 2, kx = kw / 2;
    double *dp = (double *)malloc(sizeof(double) * (size_t)ffth * fftw);
    double *res = (double *)malloc(sizeof(double) * (size_t)dh * dw);
    int x, y, i, j;
    for (y = 0; y < ffth; y++) {
        int sy = orc_pad_index(y, dh, ky);
        for (x = 0; x < fftw; x++)
            dp[(size_t)y * fftw + x] = data[sy * dw + orc_pad_index(x, dw, kx)];
    }
    for (y = 0; y < dh; y++)
        for (x = 0; x < dw; x++) {
            double acc = 0;
            for (i = 0; i < kh; i++) {
                int yy = ((y + ky - i) % ffth + ffth) % ffth;
                for (j = 0; j < kw; j++) {
                    int xx = ((x + kx - j) % fftw + fftw) % fftw;
                    acc += kernel[i * kw + j] * dp[(size_t)yy * fftw + xx];
                }
            }
            res[y * dw + x] = acc;
        }
    memcpy(data, res, sizeof(double) * (size_t)dh * dw);
    free(dp);
    free(res);
}

/* fmvd_deconvolve.cpp:139-204 (computeKernel2) in FP64, from the raw kernels.
 * psf_type: 0 independent, 1 efficient_bayesian, 2 optimization_1,
 * 3 optimization_2 (fmvd_deconvolve.h:9). */
int orc_kernel_hat(const float *kernels, int kw, int kh, int nviews,
                   int psf_type, double *out)
{
    const int len = kw * kh;
    double *raw = (double *)malloc(sizeof(double) * len * nviews);
    double *t1 = (double *)malloc(sizeof(double) * len);
    double *t2 = (double *)malloc(sizeof(double) * len);
    int v, w, i;
    if (psf_type < 0 || psf_type > 3) return -1;
    for (i = 0; i < len * nviews; i++) raw[i] = (double)kernels[i];
    for (v = 0; v < nviews; v++) {
        double *k2 = out + (size_t)len * v;
        const double *pv = raw + (size_t)len * v;
        switch (psf_type) {
        case 0:
            orc_flip_kernel_f64(pv, k2, kw, kh);
            break;
        case 1: /* :154-172  flip(Pv) * prod_{w!=v} flip(Pv)(*)Pw(*)flip(Pw) */
            orc_flip_kernel_f64(pv, k2, kw, kh);
            for (w = 0; w < nviews; w++) {
                if (w == v) continue;
                orc_flip_kernel_f64(pv, t1, kw, kh);
                orc_convolve_single_plane(t1, kw, kh, raw + (size_t)len * w, kw, kh);
                orc_flip_kernel_f64(raw + (size_t)len * w, t2, kw, kh);
                orc_convolve_single_plane(t1, kw, kh, t2, kw, kh);
                for (i = 0; i < len; i++) k2[i] *= t1[i];
            }
            break;
        case 2: /* :173-189  flip(Pv) * prod_{w!=v} flip(Pv)(*)Pw */
            orc_flip_kernel_f64(pv, k2, kw, kh);
            for (w = 0; w < nviews; w++) {
                if (w == v) continue;
                orc_flip_kernel_f64(pv, t1, kw, kh);
                orc_convolve_single_plane(t1, kw, kh, raw + (size_t)len * w, kw, kh);
                for (i = 0; i < len; i++) k2[i] *= t1[i];
            }
            break;
        case 3: /* :190-196  flip(Pv ^ V) */
            for (i = 0; i < len; i++) t1[i] = pow(pv[i], nviews);
            orc_flip_kernel_f64(t1, k2, kw, kh);
            break;
        }
        normalize_f64(k2, len); /* :198 */
    }
    free(raw);
    free(t1);
    free(t2);
    return 0;
}

void orc_normalized_weights(const float *w, int dw, int dh, int nviews,
                            float *out)
{
    const size_t dP = (size_t)dw * dh;
    size_t i;
    int v;
    for (i = 0; i < dP; i++) {
        float s = 0.0f;
        for (v = 0; v < nviews; v++) s += w[dP * v + i];
        for (v = 0; v < nviews; v++) {
            float x = w[dP * v + i];
            out[dP * v + i] = (s > 0) ? x / s : x;
        }
    }
}

/* fmvd_deconvolve_cuda.cu:13-38 (padKernel): kp[(i-kY) mod H][(j-kX) mod W] = K[i][j] */
static void pad_kernel_f64(const double *k, int kw, int kh, int ffth, int fftw,
                           double *kp)
{
    int i, j;
    memset(kp, 0, sizeof(double) * (size_t)ffth * fftw);
    for (i = 0; i < kh; i++)
        for (j = 0; j < kw; j++) {
            int y = i - kh / 2, x = j - kw / 2;
            if (y < 0) y += ffth;
            if (x < 0) x += fftw;
            kp[(size_t)y * fftw + x] = k[i * kw + j];
        }
}

int orc_build_operands_f64(int dw, int dh, int nviews, const uint16_t *plane,
                           const float *weights, float *kernels, int kw,
                           int kh, int psf_type, int ffth, int fftw,
                           double *Dp, double *Wp, double *kp, double *kph)
{
    const size_t P = (size_t)ffth * fftw, dP = (size_t)dw * dh;
    const int len = kw * kh;
    double *khat = (double *)malloc(sizeof(double) * len * nviews);
    double *kn = (double *)malloc(sizeof(double) * len);
    float *wn = (float *)malloc(sizeof(float) * dP * nviews);
    int v, x, y, i;
    /* fmvd_deconvolve.impl.cpp:61-64: kernel2 from raw kernels, THEN the
     * caller's kernels are normalised in place (float semantics). */
    if (orc_kernel_hat(kernels, kw, kh, nviews, psf_type, khat)) return -1;
    for (v = 0; v < nviews; v++) orc_normalize_f32(kernels + (size_t)len * v, len);
    orc_normalized_weights(weights, dw, dh, nviews, wn);
    for (v = 0; v < nviews; v++) {
        for (i = 0; i < len; i++) kn[i] = (double)kernels[(size_t)len * v + i];
        pad_kernel_f64(kn, kw, kh, ffth, fftw, kp + P * v);
        pad_kernel_f64(khat + (size_t)len * v, kw, kh, ffth, fftw, kph + P * v);
        for (y = 0; y < ffth; y++) {
            int sy = orc_pad_index(y, dh, kh / 2);  /* weights: natural order */
            int dy = orc_pad_index(y, dh, kw / 2);  /* data: swapped (impl.cu:76-77) */
            for (x = 0; x < fftw; x++) {
                int sx = orc_pad_index(x, dw, kw / 2);
                int dx = orc_pad_index(x, dw, kh / 2);
                Wp[P * v + (size_t)y * fftw + x] = wn[dP * v + (size_t)sy * dw + sx];
                if (plane)
                    Dp[P * v + (size_t)y * fftw + x] = plane[dP * v + (size_t)dy * dw + dx];
            }
        }
    }
    free(khat);
    free(kn);
    free(wn);
    return 0;
}

/* ---- transform (fmvd_transform.impl.cu:24-56) ----
 * Source coordinates in FP64: r = ((m0 x + m1 y) + m2 z) + m3 per axis,
 * i = floor(r), a = r - i; texels outside [0,w)x[0,h)x[0,d) read 0
 * (border address mode, texel centres at integer+0.5, :52,:123-125);
 * nested lerps x, then y, then z with a*(1-t) + b*t;
 * v = interp / (2^B - 1); out = clamp(floor(v * 2^B + 0.5)). */
static inline double tex_at(const uint16_t *src, int w, int h, int d, int x,
                            int y, int z)
{
    if (x < 0 || y < 0 || z < 0 || x >= w || y >= h || z >= d) return 0.0;
    return (double)src[((size_t)z * h + y) * w + x];
}

void orc_transform(const uint16_t *src, int w, int h, int d, int bits,
                   const float *m, int tw, int th, int z0, int nz,
                   uint16_t *out)
{
    const double maxv = (double)((1 << bits) - 1);
    const double scale = (double)(1 << bits);
    double md[12];
    int i;
    long row;
    for (i = 0; i < 12; i++) md[i] = (double)m[i];
    /* target rows are independent: OpenMP over (z, y) */
#pragma omp parallel for schedule(dynamic, 16)
    for (row = 0; row < (long)nz * th; row++) {
        const int z = z0 + (int)(row / th), y = (int)(row % th);
        int x;
            for (x = 0; x < tw; x++) {
                double rx = ((md[0] * x + md[1] * y) + md[2] * z) + md[3];
                double ry = ((md[4] * x + md[5] * y) + md[6] * z) + md[7];
                double rz = ((md[8] * x + md[9] * y) + md[10] * z) + md[11];
                double fx = floor(rx), fy = floor(ry), fz = floor(rz);
                double ax = rx - fx, ay = ry - fy, az = rz - fz;
                double bx = 1.0 - ax, by = 1.0 - ay, bz = 1.0 - az;
                double c00, c10, c01, c11, c0, c1, val, vv, yv;
                int ix, iy, iz;
                uint16_t o;
                /* coordinates far outside: every corner is outside -> 0 */
                if (!(rx > -2.0 && ry > -2.0 && rz > -2.0 && rx < (double)w + 1.0 &&
                      ry < (double)h + 1.0 && rz < (double)d + 1.0)) {
                    val = 0.0;
                } else {
                    ix = (int)fx; iy = (int)fy; iz = (int)fz;
                    c00 = tex_at(src, w, h, d, ix, iy, iz) * bx + tex_at(src, w, h, d, ix + 1, iy, iz) * ax;
                    c10 = tex_at(src, w, h, d, ix, iy + 1, iz) * bx + tex_at(src, w, h, d, ix + 1, iy + 1, iz) * ax;
                    c01 = tex_at(src, w, h, d, ix, iy, iz + 1) * bx + tex_at(src, w, h, d, ix + 1, iy, iz + 1) * ax;
                    c11 = tex_at(src, w, h, d, ix, iy + 1, iz + 1) * bx + tex_at(src, w, h, d, ix + 1, iy + 1, iz + 1) * ax;
                    c0 = c00 * by + c10 * ay;
                    c1 = c01 * by + c11 * ay;
                    val = c0 * bz + c1 * az;
                }
                vv = val / maxv;
                yv = vv * scale + 0.5;
                if (yv >= maxv) o = (uint16_t)maxv;
                else if (yv > 0) o = (uint16_t)floor(yv);
                else o = 0;
                out[((size_t)(z - z0) * th + y) * tw + x] = o;
            }
    }
}

/* fmvd_transform.cu:19-67 at z = td/2 (fmvd_transform.impl.cu:155), FP64. */
void orc_mask(int w, int h, int d, int tw, int th, int td, const float *m,
              int border, float zspacing, float *out)
{
    const int z = td / 2;
    const double b = (double)border;
    double md[12];
    int i, x, y;
    for (i = 0; i < 12; i++) md[i] = (double)m[i];
    for (y = 0; y < th; y++)
        for (x = 0; x < tw; x++) {
            double rx = ((md[0] * x + md[1] * y) + md[2] * z) + md[3];
            double ry = ((md[4] * x + md[5] * y) + md[6] * z) + md[7];
            double rz = ((md[8] * x + md[9] * y) + md[10] * z) + md[11];
            double v = 1.0, dx, dy, dz;
            if (rx < 0 || rx >= w || ry < 0 || ry >= h || rz < 0 || rz >= d) {
                out[(size_t)y * tw + x] = 0.0f;
                continue;
            }
            dx = rx < (double)(w / 2) ? rx : (double)w - rx;
            dy = ry < (double)(h / 2) ? ry : (double)h - ry;
            dz = rz < (double)(d / 2) ? rz : (double)d - rz;
            dz *= (double)zspacing;
            if (dx < b) v *= 0.5 * (1.0 - cos(dx / b * M_PI));
            if (dy < b) v *= 0.5 * (1.0 - cos(dy / b * M_PI));
            if (dz < b) v *= 0.5 * (1.0 - cos(dz / b * M_PI));
            out[(size_t)y * tw + x] = (float)v;
        }
}
