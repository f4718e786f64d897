// Host-side plan preparation (see host_psf.hpp).
#include "host_psf.hpp"

#include <cmath>
#include <complex>
#include <cstring>
#include <thread>
#include <vector>

#include "hip_check.hpp"
#include "rl_launch.hpp"

namespace fmvd {

using cd = std::complex<double>;

namespace {

int align_up(int a, int b) { return (a % b != 0) ? (a - a % b + b) : a; }

int pad_index(int i, int n, int half) { return i < n ? i : (i < n + half ? n - 1 : 0); }

// exp(-2 pi i m / n) with exact integer reduction of m
cd unit_root(long long m, long long n)
{
    m %= n;
    if (m < 0) m += n;
    const double a = -2.0 * M_PI * (double)m / (double)n;
    return cd(std::cos(a), std::sin(a));
}

// Recursive mixed-radix decimation in time.  out[k] = sum_j in[j*s] w^{jk}.
void fft_rec(cd* out, const cd* in, int n, int s, const int* fac, const std::vector<cd>& root,
             int N)
{
    if (n == 1) {
        out[0] = in[0];
        return;
    }
    const int p = fac[0], m = n / p;
    for (int q = 0; q < p; ++q) fft_rec(out + q * m, in + q * s, m, s * p, fac + 1, root, N);
    const int step = N / n;  // root[] is for length N; W_n^x = root[x*step]
    cd t[64], u[64];
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < p; ++q) t[q] = out[q * m + k] * root[(long long)q * k * step % N];
        for (int r = 0; r < p; ++r) {
            cd acc = 0;
            for (int q = 0; q < p; ++q) acc += t[q] * root[(long long)q * r * m * step % N];
            u[r] = acc;
        }
        for (int r = 0; r < p; ++r) out[k + m * r] = u[r];
    }
}

template <typename F> void parallel_for(int n, F f)
{
    unsigned nt = std::thread::hardware_concurrency();
    if (nt == 0) nt = 1;
    if (nt > 32) nt = 32;
    if ((int)nt > n) nt = (unsigned)(n > 0 ? n : 1);
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t)
        th.emplace_back([=] {
            for (int i = (int)t; i < n; i += (int)nt) f(i);
        });
    for (auto& x : th) x.join();
}

void flip(const double* in, double* out, int kw, int kh)
{
    for (int y = 0; y < kh; ++y)
        for (int x = 0; x < kw; ++x) out[y * kw + x] = in[(kh - y - 1) * kw + (kw - x - 1)];
}

// fmvd_deconvolve.cpp:90-137: clamp-padded circular convolution on the
// snap-sized torus, cropped to the data size.
void convolve_single_plane(double* data, int dw, int dh, const double* kernel, int kw, int kh)
{
    const int fh = snap_transform_size(dh + kh - 1), fw = snap_transform_size(dw + kw - 1);
    const int ky = kh / 2, kx = kw / 2;
    std::vector<double> padded((size_t)fh * fw), res((size_t)dh * dw);
    for (int y = 0; y < fh; ++y)
        for (int x = 0; x < fw; ++x)
            padded[(size_t)y * fw + x] = data[pad_index(y, dh, ky) * dw + pad_index(x, dw, kx)];
    for (int y = 0; y < dh; ++y)
        for (int x = 0; x < dw; ++x) {
            double acc = 0;
            for (int i = 0; i < kh; ++i) {
                const int yy = ((y + ky - i) % fh + fh) % fh;
                for (int j = 0; j < kw; ++j) {
                    const int xx = ((x + kx - j) % fw + fw) % fw;
                    acc += kernel[i * kw + j] * padded[(size_t)yy * fw + xx];
                }
            }
            res[(size_t)y * dw + x] = acc;
        }
    std::memcpy(data, res.data(), sizeof(double) * res.size());
}

void normalize_f64(double* k, int len)
{
    double s = 0;
    for (int i = 0; i < len; ++i) s += k[i];
    for (int i = 0; i < len; ++i) k[i] /= s;
}

}  // namespace

int snap_transform_size(int n)
{
    n = align_up(n, 16);
    int hibit;
    for (hibit = 31; hibit >= 0; --hibit)
        if ((unsigned)n & (1u << hibit)) break;
    const unsigned lo = 1u << hibit;
    if (lo == (unsigned)n) return n;
    const unsigned hi = 1u << (hibit + 1);
    if (hi <= 1024) return (int)hi;
    return align_up(n, 512);
}

void normalize_kernel_f32(float* k, int len)
{
    double sum = 0;
    for (int i = 0; i < len; ++i) sum += k[i];
    for (int i = 0; i < len; ++i) k[i] /= (float)sum;
}

std::vector<double> compute_kernel_hat(const float* const* raw, int kw, int kh, int nviews,
                                       int psf_type)
{
    const int len = kw * kh;
    std::vector<double> P((size_t)len * nviews), out((size_t)len * nviews), t1(len), t2(len);
    for (int v = 0; v < nviews; ++v)
        for (int i = 0; i < len; ++i) P[(size_t)v * len + i] = raw[v][i];
    for (int v = 0; v < nviews; ++v) {
        double* k2 = out.data() + (size_t)v * len;
        const double* pv = P.data() + (size_t)v * len;
        switch (psf_type) {
        case 0:  // independent: flip(Pv)
            flip(pv, k2, kw, kh);
            break;
        case 1:  // efficient_bayesian: flip(Pv) * prod_{w!=v} flip(Pv)(*)Pw(*)flip(Pw)
        case 2:  // optimization_1:      flip(Pv) * prod_{w!=v} flip(Pv)(*)Pw
            flip(pv, k2, kw, kh);
            for (int w = 0; w < nviews; ++w) {
                if (w == v) continue;
                const double* pw = P.data() + (size_t)w * len;
                flip(pv, t1.data(), kw, kh);
                convolve_single_plane(t1.data(), kw, kh, pw, kw, kh);
                if (psf_type == 1) {
                    flip(pw, t2.data(), kw, kh);
                    convolve_single_plane(t1.data(), kw, kh, t2.data(), kw, kh);
                }
                for (int i = 0; i < len; ++i) k2[i] *= t1[i];
            }
            break;
        case 3:  // optimization_2: flip(Pv^V)
            for (int i = 0; i < len; ++i) t1[i] = std::pow(pv[i], nviews);
            flip(t1.data(), k2, kw, kh);
            break;
        default:
            throw_msg("invalid PSF type (expected 0..3)");
        }
        normalize_f64(k2, len);
    }
    return out;
}

std::vector<float> normalized_weights(const float* const* w, int dw, int dh, int nviews)
{
    const size_t dP = (size_t)dw * dh;
    std::vector<float> out(dP * nviews);
    for (size_t i = 0; i < dP; ++i) {
        float s = 0.0f;
        for (int v = 0; v < nviews; ++v) s += w[v][i];
        for (int v = 0; v < nviews; ++v) out[(size_t)v * dP + i] = (s > 0) ? w[v][i] / s : w[v][i];
    }
    return out;
}

struct FftPlan {
    int n;
    std::vector<int> fac;
    std::vector<cd> root;
};

static FftPlan make_fft_plan(int n, int sign)
{
    FftPlan pl;
    pl.n = n;
    std::vector<int>& fac = pl.fac;
    int m = n;
    for (int p : {4, 2, 3, 5, 7})
        while (m % p == 0) {
            fac.push_back(p);
            m /= p;
        }
    for (int p = 11; m > 1; p += 2)
        while (m % p == 0) {
            fac.push_back(p);
            m /= p;
        }
    for (int f : fac)
        if (f > 64) throw_msg("host FFT: prime factor too large");
    fac.push_back(1);
    pl.root.resize(n);
    for (int k = 0; k < n; ++k) {
        cd r = unit_root(k, n);
        pl.root[k] = sign > 0 ? std::conj(r) : r;
    }
    return pl;
}

static void run_fft(const FftPlan& pl, std::vector<double>& re_im)
{
    const int n = pl.n;
    std::vector<cd> in(n), out(n);
    for (int k = 0; k < n; ++k) in[k] = cd(re_im[2 * k], re_im[2 * k + 1]);
    fft_rec(out.data(), in.data(), n, 1, pl.fac.data(), pl.root, n);
    for (int k = 0; k < n; ++k) {
        re_im[2 * k] = out[k].real();
        re_im[2 * k + 1] = out[k].imag();
    }
}

void host_fft(std::vector<double>& re_im, int n, int sign) { run_fft(make_fft_plan(n, sign), re_im); }

std::vector<float2> kernel_spectrum_transposed(const double* k, int kw, int kh, int H, int W,
                                               double scale)
{
    const int N = W / 2;
    // Row DFT of the kh non-zero rows at kx = 0..N (direct, exact phase).
    std::vector<cd> F((size_t)kh * (N + 1));
    parallel_for(kh, [&](int i) {
        for (int kx = 0; kx <= N; ++kx) {
            cd acc = 0;
            for (int j = 0; j < kw; ++j) {
                const int xj = ((j - kw / 2) % W + W) % W;
                acc += k[i * kw + j] * unit_root((long long)kx * xj, W);
            }
            F[(size_t)i * (N + 1) + kx] = acc;
        }
    });
    // Column FFT (length H) for each kx of the sparse column.
    std::vector<float2> out((size_t)(N + 1) * H);
    const FftPlan colplan = make_fft_plan(H, -1);
    parallel_for(N + 1, [&](int kx) {
        std::vector<double> col(2 * (size_t)H, 0.0);
        for (int i = 0; i < kh; ++i) {
            const int yi = ((i - kh / 2) % H + H) % H;
            col[2 * yi] += F[(size_t)i * (N + 1) + kx].real();
            col[2 * yi + 1] += F[(size_t)i * (N + 1) + kx].imag();
        }
        run_fft(colplan, col);
        // device layout (rl_launch.hpp): one contiguous column per kx
        for (int ky = 0; ky < H; ++ky)
            out[(size_t)kx * H + ky] = make_float2((float)(col[2 * ky] * scale), (float)(col[2 * ky + 1] * scale));
    });
    return out;
}

std::vector<float2> twiddles(int n, int count)
{
    std::vector<float2> t(count);
    for (int m = 0; m < count; ++m) {
        cd r = unit_root(m, n);
        t[m] = make_float2((float)r.real(), (float)r.imag());
    }
    return t;
}

}  // namespace fmvd
