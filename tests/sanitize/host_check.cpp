// Drives the library's host-side plan preparation (csrc/host_psf.cpp) under
// AddressSanitizer and UndefinedBehaviorSanitizer: FFT extents, kernel
// normalisation, the four PSF types' kernel-hat, weight normalisation with
// zero-weight pixels, kernel spectra in the device layout (small, odd and
// above-4096 extents), twiddles and the FP64 FFT.  Prints a checksum.
#include <cmath>
#include <cstdio>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "host_psf.hpp"

namespace fmvd {
[[noreturn]] void throw_msg(const std::string& msg) { throw std::runtime_error(msg); }
}  // namespace fmvd

int main()
{
    using namespace fmvd;
    std::mt19937 rng(7);
    std::uniform_real_distribution<float> u(0.0f, 1.0f);
    double sum = 0.0;
    for (int n = 1; n <= 9000; ++n) sum += snap_transform_size(n);
    for (int kh : {1, 4, 9, 16, 41})
        for (int kw : {1, 5, 12, 41})
            for (int views : {1, 2, 3})
                for (int psf = 0; psf < 4; ++psf) {
                    std::vector<std::vector<float>> k(views, std::vector<float>((size_t)kh * kw));
                    std::vector<const float*> kp;
                    for (auto& x : k) {
                        for (auto& v : x) v = u(rng) + 0.01f;
                        kp.push_back(x.data());
                    }
                    std::vector<double> hat = compute_kernel_hat(kp.data(), kw, kh, views, psf);
                    for (double v : hat) sum += v;
                    for (auto& x : k) normalize_kernel_f32(x.data(), kh * kw);
                }
    for (int dh : {1, 7, 64})
        for (int dw : {1, 6, 65}) {
            std::vector<std::vector<float>> w(3, std::vector<float>((size_t)dh * dw));
            std::vector<const float*> wp;
            for (auto& x : w) {
                for (auto& v : x) v = u(rng) < 0.3f ? 0.0f : u(rng);
                wp.push_back(x.data());
            }
            for (float v : normalized_weights(wp.data(), dw, dh, 3)) sum += v;
        }
    std::vector<double> kern(41 * 41);
    for (auto& v : kern) v = u(rng);
    for (int H : {16, 128, 1536})
        for (int W : {16, 256, 5120}) {
            if ((size_t)H * W > (size_t)1 << 20) continue;
            for (const float2& c : kernel_spectrum_transposed(kern.data(), 41, 41, H, W, 1.0 / H)) sum += c.x - c.y;
        }
    for (int n : {8, 640, 2816, 6656}) {
        for (const float2& c : twiddles(n, n / 2)) sum += c.x;
        std::vector<double> re_im(2 * (size_t)n);
        for (auto& v : re_im) v = u(rng);
        host_fft(re_im, n, -1);
        host_fft(re_im, n, +1);
        sum += re_im[0] / n;
    }
    printf("host_check ok %.6e\n", sum);
    return std::isfinite(sum) ? 0 : 1;
}
