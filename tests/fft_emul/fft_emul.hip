// Host emulator of the device FFT engine (CPU only).
//
// Runs the pass functions of gpu_deconvolution_amd/csrc/fft_device.hpp - the
// same LDS slot addressing, per-pass twiddle tables and radix DFTs the kernels
// execute - for every compiled plan, with the T threads of a sequence
// emulated one after another between LDS exchanges, and compares forward and
// inverse transforms with a naive double-precision DFT.  LDS guard zones
// filled with NaN catch any out-of-slice access.
#include <cmath>
#include <complex>
#include <cstdio>
#include <random>
#include <vector>

#include "../../gpu_deconvolution_amd/csrc/fft_device.hpp"

using namespace fmvd;
using cd = std::complex<double>;

static const int kGuard = 64;

template <typename PL> struct Emu {
    static constexpr int N = PL::N, T = PL::T, P = PL::P;
    std::vector<float2> tw = pass_twiddle_table<PL>();
    std::vector<float2> lds = std::vector<float2>(lds_len<N>() + 2 * kGuard);
    std::vector<float> ldsf = std::vector<float>(lds_len<N>() + 2 * kGuard);  // split exchanges
    std::vector<float2> v = std::vector<float2>((size_t)T * P);
    float2* L() { return lds.data() + kGuard; }
    float* LF() { return ldsf.data() + kGuard; }
    float2* V(int t) { return v.data() + (size_t)t * P; }
    bool clean = true;

    void reset_lds()
    {
        for (auto& x : lds) x = make_float2(NAN, NAN);
        for (auto& x : ldsf) x = NAN;
    }
    bool guards_clean()
    {
        for (int i = 0; i < kGuard; ++i)
            if (!std::isnan(lds[i].x) || !std::isnan(lds[lds.size() - 1 - i].x) || !std::isnan(ldsf[i]) ||
                !std::isnan(ldsf[ldsf.size() - 1 - i]))
                return false;
        return clean;
    }

    // The exchanges of fwd_exchange / inv_exchange: all T threads write, then
    // all read; split plans move the real parts, then the imaginary parts,
    // through the float buffer (the buffer is refilled with NaN before each
    // part, so a read of a slot the part did not write shows up as NaN).
    template <int I, int COMP> void fwd_part()
    {
        reset_lds();
        for (int t = 0; t < T; ++t) lds_out_slots<PL, I - 1, true, COMP>(LF(), V(t), t);
        for (int t = 0; t < T; ++t) lds_in_slots<PL, I, false, COMP>(LF(), V(t), t);
        clean = clean && guards_clean();
    }
    template <int I, int COMP> void inv_part()
    {
        reset_lds();
        for (int t = 0; t < T; ++t) lds_in_slots<PL, I + 1, true, COMP>(LF(), V(t), t);
        for (int t = 0; t < T; ++t) lds_out_slots<PL, I, false, COMP>(LF(), V(t), t);
        clean = clean && guards_clean();
    }
    template <int I> void fwd_exchange_emu()
    {
        if constexpr (PL::SPLIT) {
            fwd_part<I, 1>();
            fwd_part<I, 2>();
        } else {
            reset_lds();
            for (int t = 0; t < T; ++t) lds_out_slots<PL, I - 1, true>(L(), V(t), t);
            for (int t = 0; t < T; ++t) lds_in_slots<PL, I, false>(L(), V(t), t);
            clean = clean && guards_clean();
        }
    }
    template <int I> void inv_exchange_emu()
    {
        if constexpr (PL::SPLIT) {
            inv_part<I, 1>();
            inv_part<I, 2>();
        } else {
            reset_lds();
            for (int t = 0; t < T; ++t) lds_in_slots<PL, I + 1, true>(L(), V(t), t);
            for (int t = 0; t < T; ++t) lds_out_slots<PL, I, false>(L(), V(t), t);
            clean = clean && guards_clean();
        }
    }

    template <int I> void fwd()
    {
        if constexpr (I < PL::NP) {
            if constexpr (I > 0) {
                // as fwd_passes: the twiddles are loaded before the exchange
                std::vector<float2> w((size_t)T * tw_regs<PL, I>());
                for (int t = 0; t < T; ++t) load_twiddles<PL, I>(w.data() + (size_t)t * tw_regs<PL, I>(), t, tw.data());
                fwd_exchange_emu<I>();
                for (int t = 0; t < T; ++t) apply_twiddles<PL, I, false>(V(t), w.data() + (size_t)t * tw_regs<PL, I>());
            } else {
                for (int t = 0; t < T; ++t) pass_twiddle<PL, I, false>(V(t), t, tw.data());
            }
            for (int t = 0; t < T; ++t) pass_dft<PL, I, -1>(V(t));
            fwd<I + 1>();
        }
    }
    template <int I> void inv()
    {
        if constexpr (I >= 0) {
            // as inv_passes: small-radix passes load their twiddles before the exchange
            constexpr bool pre = I < PL::NP - 1 && PL::R::at(I) <= 4;
            std::vector<float2> w((size_t)T * tw_regs<PL, I>());
            if constexpr (pre)
                for (int t = 0; t < T; ++t) load_twiddles<PL, I>(w.data() + (size_t)t * tw_regs<PL, I>(), t, tw.data());
            if constexpr (I < PL::NP - 1) inv_exchange_emu<I>();
            for (int t = 0; t < T; ++t) {
                pass_dft<PL, I, +1>(V(t));
                if constexpr (pre) apply_twiddles<PL, I, true>(V(t), w.data() + (size_t)t * tw_regs<PL, I>());
                else pass_twiddle<PL, I, true>(V(t), t, tw.data());
            }
            inv<I - 1>();
        }
    }

    // returns max relative error of forward and inverse vs naive DFT
    double run(unsigned seed, bool& guards)
    {
        reset_lds();
        std::mt19937 rng(seed);
        std::uniform_real_distribution<double> u(-1.0, 1.0);
        std::vector<cd> x(N);
        for (auto& z : x) z = cd(u(rng), u(rng));
        constexpr int RF = first_radix<PL, -1>(), RL = last_radix<PL, -1>();
        for (int t = 0; t < T; ++t)
            for (int i = 0; i < P; ++i) {
                const cd z = x[slot_pos<PL, RF>(t, i)];
                V(t)[i] = make_float2((float)z.real(), (float)z.imag());
            }
        fwd<0>();
        guards = guards_clean();
        std::vector<cd> X(N);
        for (int t = 0; t < T; ++t)
            for (int i = 0; i < P; ++i) X[slot_pos<PL, RL>(t, i)] = cd(V(t)[i].x, V(t)[i].y);
        double err = 0, nrm = 0;
        for (int k = 0; k < N; ++k) {
            cd s = 0;
            for (int n = 0; n < N; ++n) s += x[n] * std::polar(1.0, -2.0 * M_PI * (double)((long)n * k % N) / N);
            err = std::max(err, std::abs(s - X[k]));
            nrm = std::max(nrm, std::abs(s));
        }
        double e1 = err / nrm;
        // inverse from the forward's registers: must return N * x at layout RF
        static_assert(first_radix<PL, +1>() == RL && last_radix<PL, +1>() == RF, "layout chaining");
        inv<PL::NP - 1>();
        guards = guards && guards_clean();
        err = 0;
        nrm = 0;
        for (int t = 0; t < T; ++t)
            for (int i = 0; i < P; ++i) {
                const cd got(V(t)[i].x, V(t)[i].y), want = (double)N * x[slot_pos<PL, RF>(t, i)];
                err = std::max(err, std::abs(got - want));
                nrm = std::max(nrm, std::abs(want));
            }
        return std::max(e1, err / nrm);
    }
};

static int failures = 0;

template <typename PL> void check(const char* name)
{
    static_assert(plan_valid<PL>(), "invalid plan");
    Emu<PL> e;
    bool guards = true;
    double worst = 0;
    for (unsigned s = 1; s <= 2; ++s) worst = std::max(worst, e.run(s * 7919u + PL::N, guards));
    // FP32 FFT: error grows ~ log2(N) * eps
    const double tol = 2e-6 * std::log2((double)PL::N);
    const bool ok = worst < tol && guards;
    if (!ok) ++failures;
    printf("%-26s N=%5d T=%4d P=%3d passes=%d%s relerr=%.2e guards=%s %s\n", name, PL::N, PL::T, PL::P, PL::NP,
           PL::SPLIT ? " split" : "", worst, guards ? "clean" : "TOUCHED", ok ? "ok" : "FAIL");
}

#define CHECK_ALL(n)                                                             \
    check<PlanV<n, 0>>("PlanV<" #n ",0>");                                       \
    if constexpr (PlanV<n, 1>::distinct) check<PlanV<n, 1>>("PlanV<" #n ",1>"); \
    if constexpr (PlanV<n, 2>::distinct) check<PlanV<n, 2>>("PlanV<" #n ",2>");

int main()
{
    CHECK_ALL(8) CHECK_ALL(16) CHECK_ALL(32) CHECK_ALL(64) CHECK_ALL(128) CHECK_ALL(256) CHECK_ALL(512)
    CHECK_ALL(768) CHECK_ALL(1024) CHECK_ALL(1280) CHECK_ALL(1536) CHECK_ALL(1792) CHECK_ALL(2048)
    CHECK_ALL(2560) CHECK_ALL(3072) CHECK_ALL(3584) CHECK_ALL(4096) CHECK_ALL(2304) CHECK_ALL(2816)
    CHECK_ALL(3328) CHECK_ALL(3840) CHECK_ALL(4608) CHECK_ALL(5120) CHECK_ALL(5632) CHECK_ALL(6144)
    CHECK_ALL(6656) CHECK_ALL(7168) CHECK_ALL(7680) CHECK_ALL(8192)
    printf("%s: %d plan(s) failed\n", failures ? "FAIL" : "PASS", failures);
    return failures ? 1 : 0;
}
