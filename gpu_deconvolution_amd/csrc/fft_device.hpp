// Device FFT engine for gfx950 (CDNA4): in-register small DFTs plus
// LDS-exchanged Stockham passes.
//
// A length-N complex sequence is owned by T threads of one workgroup.  Each
// thread holds P = N/T complex values in registers.  A pass of radix R makes
// every thread run P/R radix-R butterflies on the values it holds; between
// passes the values are exchanged through LDS (write outputs, barrier, read
// the next pass's inputs, barrier).  The radix plan (R0, R1, ...) is a
// compile-time list; every radix divides P.
//
// Position convention of a Stockham pass (radix R, span NS = product of the
// radices already applied, M = N/R butterflies): butterfly j reads positions
// j + r*M (r < R) and writes positions (j/NS)*NS*R + j%NS + r*NS.  Thread t
// owns butterflies j = t + b*T (b < P/R).  The input layout of the first pass
// and the output layout of the last pass are therefore both "j + r*M", which
// lets a kernel load straight from global memory into the first pass and
// store straight from the last pass, and lets an inverse transform whose first
// radix equals the forward transform's last radix start from the forward
// transform's registers with no LDS round trip.
//
// Twiddles of the Stockham passes come from a table tw[m] = exp(-2 pi i m/N)
// (m < N) computed in double on the host; the inverse conjugates them.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fmvd {

// ---------------------------------------------------------------- complex ops
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b)
{
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by (DIR * i): DIR = -1 -> -i, DIR = +1 -> +i
template <int DIR> __device__ __forceinline__ float2 mul_i(float2 a)
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
__device__ __forceinline__ constexpr float2 ct_twiddle(long m, long n)
{
    return make_float2(float(ct_cos2pi(m, n)), float(DIR * ct_sin2pi(m, n)));
}

// ------------------------------------------------------ in-register DFTs
// dft<R,DIR>(v): in-place DFT of v[0..R) with kernel exp(DIR*2*pi*i*n*k/R),
// output in natural order.  Unnormalised.
template <int R, int DIR> struct Dft;

template <int DIR> struct Dft<1, DIR> {
    static __device__ __forceinline__ void run(float2*) {}
};

template <int DIR> struct Dft<2, DIR> {
    static __device__ __forceinline__ void run(float2* v)
    {
        float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <int DIR> struct Dft<3, DIR> {
    static __device__ __forceinline__ void run(float2* v)
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
    static __device__ __forceinline__ void run(float2* v)
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
    static __device__ __forceinline__ void run(float2* v)
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
    static __device__ __forceinline__ void run(float2* v)
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

// Composite R = A*B by in-register Cooley-Tukey:
//   X[k1 + A*k2] = sum_n2 W_R^{n2 k1} [sum_n1 x[B*n1 + n2] W_A^{n1 k1}] W_B^{n2 k2}
template <int A, int B, int DIR> struct DftSplit {
    static __device__ __forceinline__ void run(float2* v)
    {
        constexpr int R = A * B;
        float2 t[R];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) {
            float2 s[A];
#pragma unroll
            for (int n1 = 0; n1 < A; ++n1) s[n1] = v[B * n1 + n2];
            Dft<A, DIR>::run(s);
#pragma unroll
            for (int k1 = 0; k1 < A; ++k1) {
                if (n2 * k1 == 0) t[k1 * B + n2] = s[k1];
                else t[k1 * B + n2] = cmul(s[k1], ct_twiddle<DIR>(n2 * k1, R));
            }
        }
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

template <int N_, int T_, typename Radices> struct PlanDef {
    static constexpr int N = N_;
    static constexpr int T = T_;
    static constexpr int P = N_ / T_;
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
template <> struct Plan<768> : PlanDef<768, 32, RadixList<12, 8, 8>> {};
template <> struct Plan<1024> : PlanDef<1024, 64, RadixList<16, 16, 4>> {};
template <> struct Plan<1280> : PlanDef<1280, 64, RadixList<20, 4, 4, 4>> {};
template <> struct Plan<1536> : PlanDef<1536, 64, RadixList<24, 8, 8>> {};
template <> struct Plan<1792> : PlanDef<1792, 64, RadixList<28, 4, 4, 4>> {};
template <> struct Plan<2048> : PlanDef<2048, 128, RadixList<16, 16, 8>> {};
template <> struct Plan<2560> : PlanDef<2560, 64, RadixList<5, 8, 8, 8>> {};
template <> struct Plan<3072> : PlanDef<3072, 64, RadixList<12, 16, 16>> {};
template <> struct Plan<3584> : PlanDef<3584, 64, RadixList<7, 8, 8, 8>> {};
template <> struct Plan<4096> : PlanDef<4096, 256, RadixList<16, 16, 16>> {};

// Alternative plans for the extents of the benchmark workloads, selected at
// run time per kernel kind (fewer registers per thread at the price of more
// LDS exchanges, or the reverse).  PlanV<N, 0> is Plan<N>.
constexpr int kPlanVariants = 3;
template <int N, int V> struct PlanV : Plan<N> {
    static constexpr bool distinct = (V == 0);
};
#define FMVD_PLAN_VARIANT(n, v, t, ...)                                          \
    template <> struct PlanV<n, v> : PlanDef<n, t, RadixList<__VA_ARGS__>> {     \
        static constexpr bool distinct = true;                                   \
    };
FMVD_PLAN_VARIANT(256, 1, 32, 8, 8, 4)
FMVD_PLAN_VARIANT(512, 1, 64, 8, 8, 8)
FMVD_PLAN_VARIANT(768, 1, 64, 12, 4, 4, 4)
FMVD_PLAN_VARIANT(768, 2, 64, 3, 4, 4, 4, 4)
FMVD_PLAN_VARIANT(1024, 1, 64, 8, 8, 16)
FMVD_PLAN_VARIANT(1024, 2, 128, 8, 8, 4, 4)
FMVD_PLAN_VARIANT(1280, 1, 32, 20, 8, 8)
FMVD_PLAN_VARIANT(1280, 2, 64, 5, 4, 4, 4, 4)
FMVD_PLAN_VARIANT(1536, 1, 64, 3, 8, 8, 8)
FMVD_PLAN_VARIANT(1536, 2, 64, 12, 4, 4, 8)
FMVD_PLAN_VARIANT(2560, 1, 64, 10, 4, 8, 8)
#undef FMVD_PLAN_VARIANT

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
// radix of forward pass i / inverse pass i (inverse runs the list reversed)
template <typename PL, int DIR> constexpr int radix_of(int i)
{
    return DIR < 0 ? PL::R::at(i) : PL::R::at(PL::NP - 1 - i);
}
template <typename PL, int DIR> constexpr int ns_of(int i)
{
    int s = 1;
    for (int k = 0; k < i; ++k) s *= radix_of<PL, DIR>(k);
    return s;
}

// LDS index with one pad slot every 16 complex values (bank spreading).
__device__ __forceinline__ int lds_pad(int i) { return i + (i >> 4); }
template <int N> constexpr int lds_len() { return N + N / 16; }

// Butterflies of one pass held in registers: v[b*R + r] is input r of
// butterfly j = t + b*T (on entry) and output r (on exit).
template <typename PL, int R, int NS, int DIR>
__device__ __forceinline__ void pass_compute(float2* v, int t, const float2* __restrict__ tw)
{
    constexpr int N = PL::N, T = PL::T, P = PL::P;
    constexpr int NB = P / R;
    constexpr int TSTEP = N / (NS * R);
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        const int j = t + b * T;
        if (NS > 1) {
            const int k = j % NS;
#pragma unroll
            for (int r = 1; r < R; ++r) {
                float2 w = tw[k * r * TSTEP];
                v[b * R + r] = DIR < 0 ? cmul(v[b * R + r], w) : cmulc(v[b * R + r], w);
            }
        }
        Dft<R, DIR>::run(v + b * R);
    }
}

// Position of input r of butterfly b for thread t in pass with radix R.
template <typename PL, int R>
__device__ __forceinline__ int in_pos(int t, int b, int r)
{
    return t + b * PL::T + r * (PL::N / R);
}
// Position of output r of butterfly b for thread t in pass (R, NS).
template <typename PL, int R, int NS>
__device__ __forceinline__ int out_pos(int t, int b, int r)
{
    const int j = t + b * PL::T;
    return (j / NS) * NS * R + (j % NS) + r * NS;
}

template <typename PL, int R, int NS>
__device__ __forceinline__ void lds_store_out(float2* lds, const float2* v, int t)
{
    constexpr int NB = PL::P / R;
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int r = 0; r < R; ++r) lds[lds_pad(out_pos<PL, R, NS>(t, b, r))] = v[b * R + r];
}
template <typename PL, int R>
__device__ __forceinline__ void lds_load_in(const float2* lds, float2* v, int t)
{
    constexpr int NB = PL::P / R;
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int r = 0; r < R; ++r) v[b * R + r] = lds[lds_pad(in_pos<PL, R>(t, b, r))];
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

// Runs passes [I, NP) of a transform whose pass I-1 has already been computed
// in registers (or I == 0 with the first pass's inputs in registers).  The
// exchange between passes goes through `lds` (this sequence's slice); all T
// threads of the sequence must call this.
template <typename PL, int DIR, int I>
__device__ __forceinline__ void run_passes_from(float2* v, float2* lds, int t, const float2* __restrict__ tw)
{
    if constexpr (I < PL::NP) {
        constexpr int R = radix_of<PL, DIR>(I);
        constexpr int NS = ns_of<PL, DIR>(I);
        if constexpr (I > 0) {
            constexpr int RP = radix_of<PL, DIR>(I - 1);
            constexpr int NSP = ns_of<PL, DIR>(I - 1);
            lds_store_out<PL, RP, NSP>(lds, v, t);
            seq_sync<PL>();
            lds_load_in<PL, R>(lds, v, t);
            seq_sync<PL>();
        }
        pass_compute<PL, R, NS, DIR>(v, t, tw);
        run_passes_from<PL, DIR, I + 1>(v, lds, t, tw);
    }
}

// Full transform, first-pass inputs already in registers at in_pos of the
// first radix; on exit the outputs are in registers at in_pos of the last
// radix of this direction (natural-order position p = j + r*N/R_last).
// The caller must have ordered any earlier LDS reads of this sequence's slice
// before the call (seq_sync or a workgroup barrier).
template <typename PL, int DIR>
__device__ __forceinline__ void fft_regs(float2* v, float2* lds, int t, const float2* __restrict__ tw)
{
    run_passes_from<PL, DIR, 0>(v, lds, t, tw);
}

// Radix of the first pass of direction DIR (defines the register layout on
// entry) and of the last pass (defines the layout on exit).
template <typename PL, int DIR> constexpr int first_radix() { return radix_of<PL, DIR>(0); }
template <typename PL, int DIR> constexpr int last_radix() { return radix_of<PL, DIR>(PL::NP - 1); }

// Natural-order position of register slot i (= b*R + r) for a layout "j + r*N/R".
template <typename PL, int R>
__device__ __forceinline__ int slot_pos(int t, int i)
{
    return in_pos<PL, R>(t, i / R, i % R);
}

}  // namespace fmvd
