// Device FFT engine for gfx950 (CDNA4): in-register small DFTs plus
// LDS-exchanged Stockham passes.
//
// A length-N complex sequence is owned by T threads of one workgroup.  Each
// thread holds P = N/T complex values in registers.  A pass of radix R makes
// every thread run P/R radix-R butterflies on the values it holds; between
// passes the values are exchanged through LDS (write outputs, sync, read the
// next pass's inputs, sync).  The radix plan (R0, R1, ...) is a compile-time
// list; every radix divides P.
//
// Forward pass i (radix R, span NS = R0*...*R(i-1), M = N/R butterflies):
// butterfly j reads positions j + r*M (r < R), multiplies input r by
// w^(k r) with w = exp(-2 pi i / (R NS)) and k = j mod NS, runs a radix-R DFT
// and writes output r to position (j/NS)*NS*R + j%NS + r*NS.  Thread t owns
// butterflies j = t + b*T (b < P/R).  The inverse transform runs the same
// passes backwards and transposed: it reads where the forward pass wrote,
// runs the inverse DFT, multiplies by the conjugate twiddles and writes where
// the forward pass read.  Hence the forward's input layout ("j + r*N/R0") is
// the inverse's output layout and the forward's output layout
// ("j + r*N/R_last") is the inverse's input layout: a kernel can load straight
// into the first pass, store straight from the last, and start an inverse
// from the forward's registers.
//
// Plans keep every pass that writes to LDS at power-of-two radix and span (a
// radix with factors 3, 5 or 7 goes last), so every LDS address and twiddle
// address of a pass is one per-thread base plus compile-time offsets; the
// affine decomposition is verified at compile time (affine_out, affine_in,
// affine_tw) and a generic path covers the rest.
//
// Twiddles come from per-pass tables computed in double on the host
// (pass_twiddle_table), stored per pass as [r-1][k] so that the lanes of a
// wavefront, which hold consecutive k, read consecutive addresses.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <type_traits>
#include <vector>

// Host-and-device inline: the pass code also runs in the host emulator that
// checks every plan against a naive DFT.
#define FMVD_HD __host__ __device__ __forceinline__

namespace fmvd {

// ---------------------------------------------------------------- complex ops
FMVD_HD float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
FMVD_HD float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
FMVD_HD float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
FMVD_HD float2 cmulc(float2 a, float2 b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
FMVD_HD float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
FMVD_HD float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by (DIR * i): DIR = -1 -> -i, DIR = +1 -> +i
template <int DIR> FMVD_HD float2 mul_i(float2 a)
{
    return DIR < 0 ? make_float2(a.y, -a.x) : make_float2(-a.y, a.x);
}

// ------------------------------------------ compile-time trigonometry (double)
// Used only for the constant twiddles inside composite in-register DFTs.
constexpr double kPi = 3.14159265358979323846264338327950288;
constexpr double ct_sin_poly(double x)
{
    // Taylor series on |x| <= pi/4, enough terms for < 1e-17.
    double x2 = x * x, term = x, sum = x;
    for (int k = 1; k < 14; ++k) {
        term *= -x2 / double((2 * k) * (2 * k + 1));
        sum += term;
    }
    return sum;
}
constexpr double ct_cos_poly(double x)
{
    double x2 = x * x, term = 1.0, sum = 1.0;
    for (int k = 1; k < 14; ++k) {
        term *= -x2 / double((2 * k - 1) * (2 * k));
        sum += term;
    }
    return sum;
}
// cos/sin of 2*pi*m/n, exact octant reduction on integers.
constexpr double ct_cos2pi(long m, long n)
{
    m %= n;
    if (m < 0) m += n;
    // reduce to first octant using symmetries on the rational angle m/n
    // angle = 2 pi m / n; use quadrant q = floor(4m/n)
    long q = (4 * m) / n;
    long r4 = 4 * m - q * n;          // remainder in units of (2pi/(4n))... angle = (pi/2)*(q + r4/n)
    double a = (kPi / 2.0) * double(r4) / double(n);  // in [0, pi/2)
    double c = 0, s = 0;
    if (a <= kPi / 4) { c = ct_cos_poly(a); s = ct_sin_poly(a); }
    else { c = ct_sin_poly(kPi / 2 - a); s = ct_cos_poly(kPi / 2 - a); }
    switch (q & 3) {
    case 0: return c;
    case 1: return -s;
    case 2: return -c;
    default: return s;
    }
}
constexpr double ct_sin2pi(long m, long n) { return ct_cos2pi(m * 4 - n, 4 * n); }

// exp(DIR * 2 pi i m / n)
template <int DIR>
FMVD_HD constexpr float2 ct_twiddle(long m, long n)
{
    return make_float2(float(ct_cos2pi(m, n)), float(DIR * ct_sin2pi(m, n)));
}

// exp(DIR 2 pi i M / N) as a compile-time constant
template <int DIR, long M, long N> FMVD_HD float2 ct_tw()
{
    constexpr float x = float(ct_cos2pi(M, N)), y = float(DIR * ct_sin2pi(M, N));
    return make_float2(x, y);
}

// Compile-time loop: f.template operator()<i>() for i = B .. E-1, so that
// constants derived from i are evaluated by the front end, not left to the
// optimiser after unrolling.
template <int B, int E, typename F> FMVD_HD void static_for(F&& f)
{
    if constexpr (B < E) {
        f.template operator()<B>();
        static_for<B + 1, E>(f);
    }
}

// ------------------------------------------------------ in-register DFTs
// dft<R,DIR>(v): in-place DFT of v[0..R) with kernel exp(DIR*2*pi*i*n*k/R),
// output in natural order.  Unnormalised.
template <int R, int DIR> struct Dft;

template <int DIR> struct Dft<1, DIR> {
    static FMVD_HD void run(float2*) {}
};

template <int DIR> struct Dft<2, DIR> {
    static FMVD_HD void run(float2* v)
    {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <int DIR> struct Dft<3, DIR> {
    static FMVD_HD void run(float2* v)
    {
        constexpr float c1 = -0.5f;
        constexpr float s1 = float(DIR) * 0.86602540378443864676f;
        float2 a = v[0], b = v[1], c = v[2];
        float2 t = cadd(b, c), d = csub(b, c);
        v[0] = cadd(a, t);
        float2 m = make_float2(a.x + c1 * t.x, a.y + c1 * t.y);
        // (s1 * i) * d = (-s1 d.y, s1 d.x)
        float2 n = make_float2(-s1 * d.y, s1 * d.x);
        v[1] = cadd(m, n);
        v[2] = csub(m, n);
    }
};

template <int DIR> struct Dft<4, DIR> {
    static FMVD_HD void run(float2* v)
    {
        float2 t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
        float2 t2 = cadd(v[1], v[3]), t3 = mul_i<DIR>(csub(v[1], v[3]));
        v[0] = cadd(t0, t2);
        v[2] = csub(t0, t2);
        v[1] = cadd(t1, t3);
        v[3] = csub(t1, t3);
    }
};

template <int DIR> struct Dft<5, DIR> {
    static FMVD_HD void run(float2* v)
    {
        constexpr float c1 = 0.30901699437494742410f;   // cos(2pi/5)
        constexpr float c2 = -0.80901699437494742410f;  // cos(4pi/5)
        constexpr float s1 = float(DIR) * 0.95105651629515357212f;
        constexpr float s2 = float(DIR) * 0.58778525229247312917f;
        float2 a = v[0];
        float2 b1 = cadd(v[1], v[4]), d1 = csub(v[1], v[4]);
        float2 b2 = cadd(v[2], v[3]), d2 = csub(v[2], v[3]);
        v[0] = make_float2(a.x + b1.x + b2.x, a.y + b1.y + b2.y);
        float2 m1 = make_float2(a.x + c1 * b1.x + c2 * b2.x, a.y + c1 * b1.y + c2 * b2.y);
        float2 m2 = make_float2(a.x + c2 * b1.x + c1 * b2.x, a.y + c2 * b1.y + c1 * b2.y);
        // i*(s1 d1 + s2 d2) and i*(s2 d1 - s1 d2)
        float2 n1 = make_float2(-(s1 * d1.y + s2 * d2.y), s1 * d1.x + s2 * d2.x);
        float2 n2 = make_float2(-(s2 * d1.y - s1 * d2.y), s2 * d1.x - s1 * d2.x);
        v[1] = cadd(m1, n1);
        v[4] = csub(m1, n1);
        v[2] = cadd(m2, n2);
        v[3] = csub(m2, n2);
    }
};

template <int DIR> struct Dft<7, DIR> {
    static FMVD_HD void run(float2* v)
    {
        constexpr float c1 = 0.62348980185873353053f, c2 = -0.22252093395631440429f,
                        c3 = -0.90096886790241912624f;
        constexpr float s1 = float(DIR) * 0.78183148246802980871f,
                        s2 = float(DIR) * 0.97492791218182360702f,
                        s3 = float(DIR) * 0.43388373911755812048f;
        float2 a = v[0];
        float2 b1 = cadd(v[1], v[6]), d1 = csub(v[1], v[6]);
        float2 b2 = cadd(v[2], v[5]), d2 = csub(v[2], v[5]);
        float2 b3 = cadd(v[3], v[4]), d3 = csub(v[3], v[4]);
        v[0] = make_float2(a.x + b1.x + b2.x + b3.x, a.y + b1.y + b2.y + b3.y);
        float2 m1 = make_float2(a.x + c1 * b1.x + c2 * b2.x + c3 * b3.x, a.y + c1 * b1.y + c2 * b2.y + c3 * b3.y);
        float2 m2 = make_float2(a.x + c2 * b1.x + c3 * b2.x + c1 * b3.x, a.y + c2 * b1.y + c3 * b2.y + c1 * b3.y);
        float2 m3 = make_float2(a.x + c3 * b1.x + c1 * b2.x + c2 * b3.x, a.y + c3 * b1.y + c1 * b2.y + c2 * b3.y);
        float e1x = s1 * d1.x + s2 * d2.x + s3 * d3.x, e1y = s1 * d1.y + s2 * d2.y + s3 * d3.y;
        float e2x = s2 * d1.x - s3 * d2.x - s1 * d3.x, e2y = s2 * d1.y - s3 * d2.y - s1 * d3.y;
        float e3x = s3 * d1.x - s1 * d2.x + s2 * d3.x, e3y = s3 * d1.y - s1 * d2.y + s2 * d3.y;
        v[1] = make_float2(m1.x - e1y, m1.y + e1x);
        v[6] = make_float2(m1.x + e1y, m1.y - e1x);
        v[2] = make_float2(m2.x - e2y, m2.y + e2x);
        v[5] = make_float2(m2.x + e2y, m2.y - e2x);
        v[3] = make_float2(m3.x - e3y, m3.y + e3x);
        v[4] = make_float2(m3.x + e3y, m3.y - e3x);
    }
};

// Odd prime R (11, 13) from the definition, folded over conjugate pairs as
// in the radix-5 and radix-7 DFTs: with b_j = v_j + v_{R-j} and
// d_j = v_j - v_{R-j} (j = 1 .. (R-1)/2),
//   X_k = v_0 + sum_j cos(2 pi j k / R) b_j + i DIR sum_j sin(2 pi j k / R) d_j,
//   X_{R-k} = the same with the sine sum negated.
// The constants come from the compile-time trigonometry above.
template <int R, int DIR> struct DftOddPrime {
    static FMVD_HD void run(float2* v)
    {
        constexpr int H = (R - 1) / 2;
        float2 b[H + 1], d[H + 1];
        const float2 a = v[0];
        float2 sum = a;
#pragma unroll
        for (int j = 1; j <= H; ++j) {
            b[j] = cadd(v[j], v[R - j]);
            d[j] = csub(v[j], v[R - j]);
            sum = cadd(sum, b[j]);
        }
        v[0] = sum;
        static_for<1, H + 1>([&]<int k>() {
            float2 m = a, e = make_float2(0.0f, 0.0f);
            static_for<1, H + 1>([&]<int j>() {
                constexpr float c = (float)ct_cos2pi((long)j * k % R, R);
                constexpr float sn = (float)(DIR * ct_sin2pi((long)j * k % R, R));
                m = make_float2(m.x + c * b[j].x, m.y + c * b[j].y);
                e = make_float2(e.x + sn * d[j].x, e.y + sn * d[j].y);
            });
            // i * e = (-e.y, e.x)
            v[k] = make_float2(m.x - e.y, m.y + e.x);
            v[R - k] = make_float2(m.x + e.y, m.y - e.x);
        });
    }
};
template <int DIR> struct Dft<11, DIR> : DftOddPrime<11, DIR> {};
template <int DIR> struct Dft<13, DIR> : DftOddPrime<13, DIR> {};

// Composite R = A*B by in-register Cooley-Tukey:
//   X[k1 + A*k2] = sum_n2 W_R^{n2 k1} [sum_n1 x[B*n1 + n2] W_A^{n1 k1}] W_B^{n2 k2}
template <int A, int B, int DIR> struct DftSplit {
    static FMVD_HD void run(float2* v)
    {
        constexpr int R = A * B;
        float2 t[R];
        static_for<0, B>([&]<int n2>() {
            float2 s[A];
#pragma unroll
            for (int n1 = 0; n1 < A; ++n1) s[n1] = v[B * n1 + n2];
            Dft<A, DIR>::run(s);
            static_for<0, A>([&]<int k1>() {
                if constexpr (n2 * k1 == 0) t[k1 * B + n2] = s[k1];
                else t[k1 * B + n2] = cmul(s[k1], ct_tw<DIR, (long)n2 * k1, (long)R>());
            });
        });
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) {
            float2 s[B];
#pragma unroll
            for (int n2 = 0; n2 < B; ++n2) s[n2] = t[k1 * B + n2];
            Dft<B, DIR>::run(s);
#pragma unroll
            for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = s[k2];
        }
    }
};

template <int DIR> struct Dft<8, DIR> : DftSplit<2, 4, DIR> {};
template <int DIR> struct Dft<6, DIR> : DftSplit<2, 3, DIR> {};
template <int DIR> struct Dft<10, DIR> : DftSplit<2, 5, DIR> {};
template <int DIR> struct Dft<12, DIR> : DftSplit<4, 3, DIR> {};
template <int DIR> struct Dft<14, DIR> : DftSplit<2, 7, DIR> {};
template <int DIR> struct Dft<16, DIR> : DftSplit<4, 4, DIR> {};
template <int DIR> struct Dft<20, DIR> : DftSplit<4, 5, DIR> {};
template <int DIR> struct Dft<24, DIR> : DftSplit<8, 3, DIR> {};
template <int DIR> struct Dft<28, DIR> : DftSplit<4, 7, DIR> {};
template <int DIR> struct Dft<32, DIR> : DftSplit<4, 8, DIR> {};
template <int DIR> struct Dft<40, DIR> : DftSplit<8, 5, DIR> {};
template <int DIR> struct Dft<48, DIR> : DftSplit<16, 3, DIR> {};
template <int DIR> struct Dft<56, DIR> : DftSplit<8, 7, DIR> {};
template <int DIR> struct Dft<9, DIR> : DftSplit<3, 3, DIR> {};
template <int DIR> struct Dft<15, DIR> : DftSplit<3, 5, DIR> {};

// --------------------------------------------------------------- FFT plans
// Plan<N>: radices R[0..NP) (product N), threads per sequence T, P = N/T.
template <int... Rs> struct RadixList {
    static constexpr int count = sizeof...(Rs);
    static constexpr int at(int i)
    {
        constexpr int a[] = {Rs...};
        return a[i];
    }
};

// G_ (optional): rows per workgroup of the row kernels, sequences per
// workgroup of the column kernel (0 = their default).
// W_ (optional): minimum resident waves per SIMD the kernels are compiled for
// (amdgpu_waves_per_eu; 1 = whatever the workgroup size implies); a higher
// value caps the VGPRs per thread so that more workgroups fit on a CU.
// S_ (optional): split exchanges.  Every LDS exchange moves the real parts,
// then the imaginary parts, through a float buffer (E = float) instead of
// moving complex values through a float2 buffer (E = float2): half the LDS per
// sequence, for twice the synchronisations.
template <int N_, int T_, typename Radices, int G_ = 0, int W_ = 1, bool S_ = false> struct PlanDef {
    static constexpr int N = N_;
    static constexpr int T = T_;
    static constexpr int G = G_;
    static constexpr int WPE = W_;
    static constexpr int P = N_ / T_;
    static constexpr bool SPLIT = S_;
    using E = std::conditional_t<S_, float, float2>;
    using R = Radices;
    static constexpr int NP = Radices::count;
};

// Default plans.  Constraint: every radix divides P.
template <int N> struct Plan;
template <> struct Plan<8> : PlanDef<8, 1, RadixList<8>> {};
template <> struct Plan<16> : PlanDef<16, 1, RadixList<16>> {};
template <> struct Plan<32> : PlanDef<32, 2, RadixList<16, 2>> {};
template <> struct Plan<64> : PlanDef<64, 4, RadixList<16, 4>> {};
template <> struct Plan<128> : PlanDef<128, 8, RadixList<16, 8>> {};
template <> struct Plan<256> : PlanDef<256, 16, RadixList<16, 16>> {};
template <> struct Plan<512> : PlanDef<512, 32, RadixList<16, 16, 2>> {};
template <> struct Plan<768> : PlanDef<768, 64, RadixList<4, 4, 4, 12>, 8, 4> {};
template <> struct Plan<1024> : PlanDef<1024, 64, RadixList<16, 16, 4>> {};
template <> struct Plan<1280> : PlanDef<1280, 64, RadixList<4, 4, 4, 20>, 8, 4, true> {};
template <> struct Plan<1536> : PlanDef<1536, 64, RadixList<8, 8, 24>> {};
template <> struct Plan<1792> : PlanDef<1792, 64, RadixList<4, 4, 4, 28>> {};
template <> struct Plan<2048> : PlanDef<2048, 128, RadixList<16, 16, 8>> {};
template <> struct Plan<2560> : PlanDef<2560, 64, RadixList<8, 8, 8, 5>> {};
template <> struct Plan<3072> : PlanDef<3072, 64, RadixList<16, 16, 12>> {};
template <> struct Plan<3584> : PlanDef<3584, 64, RadixList<8, 8, 8, 7>> {};
template <> struct Plan<4096> : PlanDef<4096, 256, RadixList<16, 16, 16>> {};
// Extents above 4096 (snapTransformSize gives multiples of 512 there): row
// lengths N = fftW / 2 of 2304 ... 3840 and column lengths 4608 ... 8192.
// Radices 9, 11, 13 and 15 go last; every radix divides P.
template <> struct Plan<2304> : PlanDef<2304, 64, RadixList<4, 4, 4, 4, 9>> {};
template <> struct Plan<2816> : PlanDef<2816, 64, RadixList<4, 4, 4, 4, 11>> {};
template <> struct Plan<3328> : PlanDef<3328, 64, RadixList<4, 4, 4, 4, 13>> {};
template <> struct Plan<3840> : PlanDef<3840, 64, RadixList<4, 4, 4, 4, 15>> {};
template <> struct Plan<4608> : PlanDef<4608, 128, RadixList<4, 4, 4, 4, 2, 9>> {};
template <> struct Plan<5120> : PlanDef<5120, 256, RadixList<4, 4, 4, 4, 4, 5>> {};
template <> struct Plan<5632> : PlanDef<5632, 128, RadixList<4, 4, 4, 4, 2, 11>> {};
template <> struct Plan<6144> : PlanDef<6144, 256, RadixList<8, 8, 8, 12>> {};
template <> struct Plan<6656> : PlanDef<6656, 128, RadixList<4, 4, 4, 4, 2, 13>> {};
template <> struct Plan<7168> : PlanDef<7168, 256, RadixList<4, 4, 4, 4, 4, 7>> {};
template <> struct Plan<7680> : PlanDef<7680, 128, RadixList<4, 4, 4, 4, 2, 15>> {};
template <> struct Plan<8192> : PlanDef<8192, 256, RadixList<16, 16, 32>> {};

// Alternative plans for the extents of the benchmark workloads, selected at
// run time per kernel kind (fewer registers per thread at the price of more
// LDS exchanges, or the reverse).  PlanV<N, 0> is Plan<N>.
constexpr int kPlanVariants = 3;
template <int N, int V> struct PlanV : Plan<N> {
    static constexpr bool distinct = (V == 0);
    static constexpr int variant = 0;
};
#define FMVD_PLAN_VARIANT_GWS(n, v, t, g, w, s, ...)                                  \
    template <> struct PlanV<n, v> : PlanDef<n, t, RadixList<__VA_ARGS__>, g, w, s> {  \
        static constexpr bool distinct = true;                                        \
        static constexpr int variant = v;                                             \
    };
#define FMVD_PLAN_VARIANT_GW(n, v, t, g, w, ...) FMVD_PLAN_VARIANT_GWS(n, v, t, g, w, false, __VA_ARGS__)
#define FMVD_PLAN_VARIANT_G(n, v, t, g, ...) FMVD_PLAN_VARIANT_GW(n, v, t, g, 1, __VA_ARGS__)
#define FMVD_PLAN_VARIANT(n, v, t, ...) FMVD_PLAN_VARIANT_G(n, v, t, 0, __VA_ARGS__)
FMVD_PLAN_VARIANT(256, 1, 32, 8, 8, 4)
FMVD_PLAN_VARIANT(512, 1, 64, 8, 8, 8)
FMVD_PLAN_VARIANT_GW(768, 1, 64, 4, 5, 4, 4, 4, 12)
FMVD_PLAN_VARIANT_GW(768, 2, 64, 4, 6, 4, 4, 4, 12)
FMVD_PLAN_VARIANT(1024, 1, 64, 8, 8, 16)
FMVD_PLAN_VARIANT(1024, 2, 128, 8, 8, 4, 4)
FMVD_PLAN_VARIANT_GW(1280, 1, 64, 4, 4, 4, 4, 4, 20)
// variant 1 with split exchanges: 22 KB of LDS per 4-row workgroup, so four
// workgroups (4 waves per SIMD) fit where variant 1's 44 KB allow three
FMVD_PLAN_VARIANT_GWS(1280, 2, 64, 4, 4, true, 4, 4, 4, 20)
FMVD_PLAN_VARIANT(1536, 1, 64, 8, 8, 8, 3)
FMVD_PLAN_VARIANT(1536, 2, 128, 4, 4, 4, 4, 6)
// one 2-wavefront sequence per column workgroup (22 KB of LDS, 4 waves per
// SIMD), radix 20 last: 12 % faster than variant 2 on the 2560-point columns
FMVD_PLAN_VARIANT_GW(2560, 1, 128, 1, 4, 2, 4, 4, 4, 20)
FMVD_PLAN_VARIANT(2560, 2, 128, 4, 4, 4, 4, 10)
#undef FMVD_PLAN_VARIANT
#undef FMVD_PLAN_VARIANT_G
#undef FMVD_PLAN_VARIANT_GW
#undef FMVD_PLAN_VARIANT_GWS

// A plan is valid when its radices multiply to N and each divides P.
template <typename PL> constexpr bool plan_valid()
{
    int prod = 1;
    for (int i = 0; i < PL::NP; ++i) {
        if (PL::P % PL::R::at(i) != 0) return false;
        prod *= PL::R::at(i);
    }
    return prod == PL::N && PL::N % PL::T == 0;
}

// Compile-time product of radices [0, i)
template <typename PL> constexpr int span_before(int i)
{
    int s = 1;
    for (int k = 0; k < i; ++k) s *= PL::R::at(k);
    return s;
}
// Radix of step i of direction DIR, in the order the steps run: the forward
// runs passes 0..NP-1, the inverse runs them NP-1..0.
template <typename PL, int DIR> constexpr int radix_of(int i)
{
    return DIR < 0 ? PL::R::at(i) : PL::R::at(PL::NP - 1 - i);
}
// Radix whose "j + r*N/R" layout a direction consumes / produces.
template <typename PL, int DIR> constexpr int first_radix() { return radix_of<PL, DIR>(0); }
template <typename PL, int DIR> constexpr int last_radix() { return radix_of<PL, DIR>(PL::NP - 1); }

// Offset of forward pass i's twiddles in the plan's table (pass 0 has NS = 1
// and none); the inverse uses the same entries conjugated.
template <typename PL> constexpr int tw_pass_offset(int i)
{
    int off = 0;
    for (int q = 1; q < i; ++q) off += (PL::R::at(q) - 1) * span_before<PL>(q);
    return off;
}
template <typename PL> constexpr int tw_table_len() { return tw_pass_offset<PL>(PL::NP); }

// Host: the per-pass twiddle table of plan PL.
template <typename PL> std::vector<float2> pass_twiddle_table()
{
    std::vector<float2> t((size_t)tw_table_len<PL>());
    for (int q = 1; q < PL::NP; ++q) {
        const int R = PL::R::at(q), NS = span_before<PL>(q), off = tw_pass_offset<PL>(q);
        const double m = (double)R * NS;
        for (int r = 1; r < R; ++r)
            for (int k = 0; k < NS; ++k) {
                // exact argument reduction on the integer k*r mod (R*NS)
                const long e = ((long)k * r) % (long)(R * NS);
                const double a = -2.0 * 3.14159265358979323846 * (double)e / m;
                t[(size_t)off + (size_t)(r - 1) * NS + k] = make_float2((float)std::cos(a), (float)std::sin(a));
            }
    }
    return t;
}

// LDS index with one pad slot every 16 complex values (bank spreading).
FMVD_HD constexpr int lds_pad(int i) { return i + (i >> 4); }
template <int N> constexpr int lds_len() { return N + N / 16; }

// Positions of forward pass I (see the header comment).
template <typename PL, int I> FMVD_HD constexpr int in_pos_p(int t, int b, int r)
{
    return t + b * PL::T + r * (PL::N / PL::R::at(I));
}
template <typename PL, int I> FMVD_HD constexpr int out_pos_p(int t, int b, int r)
{
    constexpr int R = PL::R::at(I), NS = span_before<PL>(I);
    const int j = t + b * PL::T;
    return (j / NS) * NS * R + j % NS + r * NS;
}
template <typename PL, int I> FMVD_HD constexpr int tw_k(int t, int b)
{
    return (t + b * PL::T) % span_before<PL>(I);
}

// Affine decompositions, verified exhaustively at compile time:
//   lds_pad(pos(t, b, r)) == lds_pad(pos(t, 0, 0)) + lds_pad(pos(0, b, r))
//   tw_k(t, b) == tw_k(t, 0) + tw_k(0, b)
template <typename PL, int I> constexpr bool affine_out()
{
    for (int t = 0; t < PL::T; ++t)
        for (int b = 0; b < PL::P / PL::R::at(I); ++b)
            for (int r = 0; r < PL::R::at(I); ++r)
                if (lds_pad(out_pos_p<PL, I>(t, b, r)) !=
                    lds_pad(out_pos_p<PL, I>(t, 0, 0)) + lds_pad(out_pos_p<PL, I>(0, b, r)))
                    return false;
    return true;
}
template <typename PL, int I> constexpr bool affine_in()
{
    for (int t = 0; t < PL::T; ++t)
        for (int b = 0; b < PL::P / PL::R::at(I); ++b)
            for (int r = 0; r < PL::R::at(I); ++r)
                if (lds_pad(in_pos_p<PL, I>(t, b, r)) != lds_pad(t) + lds_pad(in_pos_p<PL, I>(0, b, r))) return false;
    return true;
}
template <typename PL, int I> constexpr bool affine_tw()
{
    for (int t = 0; t < PL::T; ++t)
        for (int b = 0; b < PL::P / PL::R::at(I); ++b)
            if (tw_k<PL, I>(t, b) != tw_k<PL, I>(t, 0) + tw_k<PL, I>(0, b)) return false;
    return true;
}

// Component COMP of a complex value as one LDS element: 0 the complex value
// (float2), 1 its real part, 2 its imaginary part (float).
template <int COMP> struct Comp;
template <> struct Comp<0> {
    using E = float2;
    static FMVD_HD float2& of(float2& v) { return v; }
};
template <> struct Comp<1> {
    using E = float;
    static FMVD_HD float& of(float2& v) { return v.x; }
};
template <> struct Comp<2> {
    using E = float;
    static FMVD_HD float& of(float2& v) { return v.y; }
};

// LDS traffic of pass I.  "in" slots are the positions the forward reads
// (and the inverse writes), "out" slots those the forward writes (and the
// inverse reads).  v[b*R + r] is slot (b, r).  COMP selects what moves: the
// complex values through a float2 buffer, or one part through a float buffer.
template <typename PL, int I, bool STORE, int COMP = 0>
FMVD_HD void lds_in_slots(typename Comp<COMP>::E* lds, float2* v, int t)
{
    using X = Comp<COMP>;
    constexpr int R = PL::R::at(I), NB = PL::P / R;
    if constexpr (affine_in<PL, I>()) {
        typename X::E* base = lds + lds_pad(t);
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                typename X::E& x = base[lds_pad(in_pos_p<PL, I>(0, b, r))];
                if constexpr (STORE) x = X::of(v[b * R + r]);
                else X::of(v[b * R + r]) = x;
            }
    } else {
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                typename X::E& x = lds[lds_pad(in_pos_p<PL, I>(t, b, r))];
                if constexpr (STORE) x = X::of(v[b * R + r]);
                else X::of(v[b * R + r]) = x;
            }
    }
}
template <typename PL, int I, bool STORE, int COMP = 0>
FMVD_HD void lds_out_slots(typename Comp<COMP>::E* lds, float2* v, int t)
{
    using X = Comp<COMP>;
    constexpr int R = PL::R::at(I), NB = PL::P / R;
    if constexpr (affine_out<PL, I>()) {
        typename X::E* base = lds + lds_pad(out_pos_p<PL, I>(t, 0, 0));
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                typename X::E& x = base[lds_pad(out_pos_p<PL, I>(0, b, r))];
                if constexpr (STORE) x = X::of(v[b * R + r]);
                else X::of(v[b * R + r]) = x;
            }
    } else {
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                typename X::E& x = lds[lds_pad(out_pos_p<PL, I>(t, b, r))];
                if constexpr (STORE) x = X::of(v[b * R + r]);
                else X::of(v[b * R + r]) = x;
            }
    }
}

// Twiddles of pass I: forward multiplies input r by w^(k r), the inverse
// multiplies output r of its inverse DFT by the conjugate;
// w = exp(-2 pi i / (R NS)), k = tw_k(t, b).
//
// Where the twiddle addresses are affine (every plan's pass that has
// twiddles), they are generated rather than loaded one by one: each thread
// loads w^k and, for radices above 8, w^(Q k) with Q = ceil(sqrt(R)) from the
// pass's table, and forms w^(k r) = w^(k a) w^(Q k m) (r = a + Q m) with the
// powers built by binary splitting: for the radices up to 36 that the plans
// use, at most four rounded products deep.
// Butterflies whose k differ (k_b = k_0 + d_b, d_b a compile-time offset)
// load their own pair of powers.  One or two loads per distinct k replace
// NB (R - 1): the loads' data return, not the arithmetic, bounds the passes
// on gfx950.
template <int R> constexpr int tw_gen_q()
{
    if (R <= 8) return R;
    int q = 1;
    while (q * q < R) ++q;
    return q;
}
template <typename PL, int I> constexpr bool tw_generated() { return span_before<PL>(I) > 1 && affine_tw<PL, I>(); }
// Whether the butterflies of a generated pass share one k (d_b = 0 for all b).
template <typename PL, int I> constexpr bool tw_shared_k()
{
    for (int b = 0; b < PL::P / PL::R::at(I); ++b)
        if (tw_k<PL, I>(0, b) != 0) return false;
    return true;
}

// Registers a pass's twiddles occupy between their load and their use.
template <typename PL, int I> constexpr int tw_regs()
{
    if constexpr (tw_generated<PL, I>()) return tw_shared_k<PL, I>() ? 2 : 2 * (PL::P / PL::R::at(I));
    else return span_before<PL>(I) > 1 ? (PL::P / PL::R::at(I)) * (PL::R::at(I) - 1) : 1;
}

// Load a pass's twiddles (generated: their base powers) into w.
template <typename PL, int I>
FMVD_HD void load_twiddles(float2* w, int t, const float2* __restrict__ tw)
{
    constexpr int R = PL::R::at(I), NB = PL::P / R, NS = span_before<PL>(I);
    if constexpr (NS > 1) {
        const float2* twp = tw + tw_pass_offset<PL>(I);
        if constexpr (tw_generated<PL, I>()) {
            constexpr int Q = tw_gen_q<R>(), NK = tw_shared_k<PL, I>() ? 1 : NB;
            const float2* tb = twp + tw_k<PL, I>(t, 0);
            static_for<0, NK>([&]<int b>() {
                constexpr int d = tw_k<PL, I>(0, b);
                w[2 * b] = tb[d];                                    // w^k (r = 1)
                if constexpr (Q < R) w[2 * b + 1] = tb[d + (Q - 1) * NS];  // w^(Q k) (r = Q)
            });
        } else {
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                const float2* tb = twp + tw_k<PL, I>(t, b);
#pragma unroll
                for (int r = 1; r < R; ++r) w[b * (R - 1) + r - 1] = tb[(r - 1) * NS];
            }
        }
    }
}

// Multiply by the twiddles load_twiddles brought into w (CONJ: their conjugates).
template <typename PL, int I, bool CONJ> FMVD_HD void apply_twiddles(float2* v, const float2* w)
{
    constexpr int R = PL::R::at(I), NB = PL::P / R, NS = span_before<PL>(I);
    if constexpr (NS > 1) {
        if constexpr (tw_generated<PL, I>()) {
            constexpr int Q = tw_gen_q<R>(), M = (R + Q - 1) / Q;
            constexpr bool shared = tw_shared_k<PL, I>();
            float2 pa[Q], pm[M > 1 ? M : 1];
            static_for<0, NB>([&]<int b>() {
                if constexpr (b == 0 || !shared) {
                    pa[1] = w[2 * b];
#pragma unroll
                    for (int n = 2; n < Q; ++n) pa[n] = cmul(pa[n / 2], pa[n - n / 2]);
                    if constexpr (M > 1) {
                        pm[1] = w[2 * b + 1];
#pragma unroll
                        for (int n = 2; n < M; ++n) pm[n] = cmul(pm[n / 2], pm[n - n / 2]);
                    }
                }
                static_for<1, R>([&]<int r>() {
                    constexpr int a = r % Q, m = r / Q;
                    float2 x;
                    if constexpr (a == 0) x = pm[m];
                    else if constexpr (m == 0) x = pa[a];
                    else x = cmul(pa[a], pm[m]);
                    v[b * R + r] = CONJ ? cmulc(v[b * R + r], x) : cmul(v[b * R + r], x);
                });
            });
        } else {
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const float2 x = w[b * (R - 1) + r - 1];
                    v[b * R + r] = CONJ ? cmulc(v[b * R + r], x) : cmul(v[b * R + r], x);
                }
        }
    }
}

// Load and apply at once.
template <typename PL, int I, bool CONJ>
FMVD_HD void pass_twiddle(float2* v, int t, const float2* __restrict__ tw)
{
    if constexpr (span_before<PL>(I) > 1) {
        float2 w[tw_regs<PL, I>()];
        load_twiddles<PL, I>(w, t, tw);
        apply_twiddles<PL, I, CONJ>(v, w);
    }
}

template <typename PL, int I, int DIR> FMVD_HD void pass_dft(float2* v)
{
    constexpr int R = PL::R::at(I), NB = PL::P / R;
#pragma unroll
    for (int b = 0; b < NB; ++b) Dft<R, DIR>::run(v + b * R);
}

// Synchronisation of the threads that own one sequence.  When a sequence's T
// threads lie inside one wavefront (T <= 64, sequences aligned to T), LDS
// traffic between them only needs wavefront ordering: a wavefront's LDS
// instructions execute in issue order, so a compiler barrier plus wavefront
// fences suffice and other wavefronts of the workgroup keep running.
// Larger sequences need the workgroup barrier.
template <typename PL> __device__ __forceinline__ void seq_sync()
{
    if constexpr (PL::T <= 64 && (64 % PL::T) == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

// The exchange before forward pass I: pass I-1's outputs go to their "out"
// slots, pass I's inputs come from their "in" slots; for split plans the real
// parts, then the imaginary parts.  Each read is ordered after every write
// of its round and before the next round's writes.
template <typename PL, int I>
__device__ __forceinline__ void fwd_exchange(typename PL::E* lds, float2* v, int t)
{
    if constexpr (PL::SPLIT) {
        lds_out_slots<PL, I - 1, true, 1>(lds, v, t);
        seq_sync<PL>();
        lds_in_slots<PL, I, false, 1>(lds, v, t);
        seq_sync<PL>();
        lds_out_slots<PL, I - 1, true, 2>(lds, v, t);
        seq_sync<PL>();
        lds_in_slots<PL, I, false, 2>(lds, v, t);
        seq_sync<PL>();
    } else {
        lds_out_slots<PL, I - 1, true>(lds, v, t);
        seq_sync<PL>();
        lds_in_slots<PL, I, false>(lds, v, t);
        seq_sync<PL>();
    }
}
// The exchange before inverse pass I (transposed: pass I+1's "in" slots to
// pass I's "out" slots).
template <typename PL, int I>
__device__ __forceinline__ void inv_exchange(typename PL::E* lds, float2* v, int t)
{
    if constexpr (PL::SPLIT) {
        lds_in_slots<PL, I + 1, true, 1>(lds, v, t);
        seq_sync<PL>();
        lds_out_slots<PL, I, false, 1>(lds, v, t);
        seq_sync<PL>();
        lds_in_slots<PL, I + 1, true, 2>(lds, v, t);
        seq_sync<PL>();
        lds_out_slots<PL, I, false, 2>(lds, v, t);
        seq_sync<PL>();
    } else {
        lds_in_slots<PL, I + 1, true>(lds, v, t);
        seq_sync<PL>();
        lds_out_slots<PL, I, false>(lds, v, t);
        seq_sync<PL>();
    }
}

// Where a pass's twiddles come from: the plan's per-pass table, read as
// each pass needs it.
struct TwTable {
    const float2* __restrict__ tw;
};
template <typename PL, int I> FMVD_HD void tw_get(float2* w, int t, const TwTable& src)
{
    load_twiddles<PL, I>(w, t, src.tw);
}

// Forward passes I..NP-1; on entry v holds pass I-1's outputs (or, for I = 0,
// the inputs at the "j + r*N/R0" layout).
template <typename PL, int I, typename TS>
__device__ __forceinline__ void fwd_passes(float2* v, typename PL::E* lds, int t, const TS& ts)
{
    if constexpr (I < PL::NP) {
        // this pass's twiddles load while the exchange runs
        float2 w[tw_regs<PL, I>()];
        tw_get<PL, I>(w, t, ts);
        if constexpr (I > 0) fwd_exchange<PL, I>(lds, v, t);
        apply_twiddles<PL, I, false>(v, w);
        pass_dft<PL, I, -1>(v);
        fwd_passes<PL, I + 1>(v, lds, t, ts);
    }
}
// Inverse passes I..0 (transposed forward passes, run backwards); on entry v
// holds the values at pass I's "out" slots.  Small-radix passes load their
// twiddles before the exchange (they are applied after the DFT, whose few
// temporaries leave room); larger ones after the DFT.
template <typename PL, int I, typename TS>
__device__ __forceinline__ void inv_passes(float2* v, typename PL::E* lds, int t, const TS& ts)
{
    if constexpr (I >= 0) {
        constexpr bool early = I < PL::NP - 1 && PL::R::at(I) <= 4;
        float2 w[tw_regs<PL, I>()];
        if constexpr (early) tw_get<PL, I>(w, t, ts);
        if constexpr (I < PL::NP - 1) inv_exchange<PL, I>(lds, v, t);
        pass_dft<PL, I, +1>(v);
        if constexpr (!early) tw_get<PL, I>(w, t, ts);
        apply_twiddles<PL, I, true>(v, w);
        inv_passes<PL, I - 1>(v, lds, t, ts);
    }
}

// Full unnormalised transform of the sequence owned by T threads, kernel
// exp(DIR 2 pi i n k / N).  On entry v holds the inputs at the layout
// "j + r*N/R" of R = first_radix<PL, DIR>() (slot_pos); on exit the outputs,
// in natural order, at the layout of last_radix<PL, DIR>().  The caller must
// have ordered any earlier LDS reads of this sequence's slice before the
// call (seq_sync or a workgroup barrier).  tw: the plan's twiddle table.
template <typename PL, int DIR>
__device__ __forceinline__ void fft_regs(float2* v, typename PL::E* lds, int t, const float2* __restrict__ tw)
{
    if constexpr (DIR < 0) fwd_passes<PL, 0>(v, lds, t, TwTable{tw});
    else inv_passes<PL, PL::NP - 1>(v, lds, t, TwTable{tw});
}

// Position of input r of butterfly b for thread t in the layout of radix R.
template <typename PL, int R> FMVD_HD int in_pos(int t, int b, int r) { return t + b * PL::T + r * (PL::N / R); }

// Natural-order position of register slot i (= b*R + r) for a layout "j + r*N/R".
template <typename PL, int R> FMVD_HD int slot_pos(int t, int i) { return in_pos<PL, R>(t, i / R, i % R); }

}  // namespace fmvd
