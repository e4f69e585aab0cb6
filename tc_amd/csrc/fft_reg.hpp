// fft_reg.hpp -- register-resident mixed-radix FFTs for gfx950.
//
// One thread owns one whole 1D line (a z-pencil, an x-row or a y-column of
// the FFT grid) in VGPRs and transforms it with a fully unrolled prime-factor
// split (power-of-two part x odd part, no twiddles between them) over Stockham
// autosort sequences of radix-8/5/4/3/2 butterflies.  Every index and twiddle
// factor is a compile-time constant: the twiddles are evaluated in double by
// constexpr code (exact integer range reduction + Taylor series) and rounded
// once to the working precision, so the FP32 path carries correctly rounded
// twiddles.  Lengths are the grid sizes of plane-wave cells (products of 2, 3
// and 5 up to 64).
//
// Sign convention (reference src/plane_wave_basis.cpp:189-196, 274-289):
//   SIGN = -1 : sum_x f(x) exp(-2 pi i k x / n)   (FFT_forward, 1/N applied by caller)
//   SIGN = +1 : sum_k c(k) exp(+2 pi i k x / n)   (FFT_backward, unscaled)
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>
#include <utility>

namespace tcx {

// ---------------------------------------------------------------------------
// complex arithmetic
// ---------------------------------------------------------------------------
template <typename T>
struct alignas(2 * sizeof(T)) cpx {
    T re, im;
};

template <typename T>
__host__ __device__ __forceinline__ cpx<T> make_cpx(T r, T i) { return cpx<T>{r, i}; }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator+(cpx<T> a, cpx<T> b) { return {a.re + b.re, a.im + b.im}; }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator-(cpx<T> a, cpx<T> b) { return {a.re - b.re, a.im - b.im}; }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> operator*(cpx<T> a, cpx<T> b)
{
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T>
__host__ __device__ __forceinline__ cpx<T> scale(cpx<T> a, T s) { return {a.re * s, a.im * s}; }
template <typename T>
__host__ __device__ __forceinline__ cpx<T> conjg(cpx<T> a) { return {a.re, -a.im}; }
// conj(a) * b
template <typename T>
__host__ __device__ __forceinline__ cpx<T> cmulc(cpx<T> a, cpx<T> b)
{
    return {a.re * b.re + a.im * b.im, a.re * b.im - a.im * b.re};
}
// a * (i * sign * s)
template <typename T, int SIGN>
__host__ __device__ __forceinline__ cpx<T> mul_isign(cpx<T> a, T s)
{
    return SIGN > 0 ? cpx<T>{-a.im * s, a.re * s} : cpx<T>{a.im * s, -a.re * s};
}

// ---------------------------------------------------------------------------
// constexpr trigonometry (double), exact for the rational angles 2*pi*m/n
// ---------------------------------------------------------------------------
constexpr double kPi = 3.14159265358979323846;

constexpr double ce_sin_poly(double x)  // |x| <= pi/4
{
    double x2 = x * x, term = x, sum = x;
    for (int k = 1; k <= 13; ++k) {
        term *= -x2 / double((2 * k) * (2 * k + 1));
        sum += term;
    }
    return sum;
}
constexpr double ce_cos_poly(double x)  // |x| <= pi/4
{
    double x2 = x * x, term = 1.0, sum = 1.0;
    for (int k = 1; k <= 13; ++k) {
        term *= -x2 / double((2 * k - 1) * (2 * k));
        sum += term;
    }
    return sum;
}

struct ce_cs {
    double c, s;
};

// (cos, sin) of 2*pi*m/n with exact integer reduction to [0, pi/4]
constexpr ce_cs ce_cis(long m, long n)
{
    m %= n;
    if (m < 0) m += n;
    const long q = (4 * m) / n;      // quadrant
    const long r = 4 * m - q * n;    // angle within quadrant = (pi/2) r / n
    double c = 0, s = 0;
    if (2 * r <= n) {
        const double x = (kPi / 2) * double(r) / double(n);
        c = ce_cos_poly(x);
        s = ce_sin_poly(x);
    } else {
        const double x = (kPi / 2) * double(n - r) / double(n);
        c = ce_sin_poly(x);
        s = ce_cos_poly(x);
    }
    switch (q) {
        case 0: return {c, s};
        case 1: return {-s, c};
        case 2: return {-c, -s};
        default: return {s, -c};
    }
}

// W_n^m = exp(SIGN * 2 pi i m / n) in precision T
template <typename T, int N, int SIGN, int M>
__host__ __device__ __forceinline__ constexpr cpx<T> twiddle()
{
    constexpr ce_cs v = ce_cis(M, N);
    return cpx<T>{static_cast<T>(v.c), static_cast<T>(SIGN * v.s)};
}

// ---------------------------------------------------------------------------
// compile-time loop
// ---------------------------------------------------------------------------
template <int B, int E, typename F>
__host__ __device__ __forceinline__ void static_for(F&& f)
{
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

// ---------------------------------------------------------------------------
// small DFT kernels:  v[s] <- sum_r v[r] W_R^{r s}
// ---------------------------------------------------------------------------
template <typename T, int R, int SIGN>
struct Dft {
    __host__ __device__ __forceinline__ static void run(cpx<T> (&v)[R])
    {
        cpx<T> y[R];
        static_for<0, R>([&](auto sc) {
            constexpr int s = decltype(sc)::value;
            cpx<T> acc = v[0];
            static_for<1, R>([&](auto rc) {
                constexpr int r = decltype(rc)::value;
                acc = acc + v[r] * twiddle<T, R, SIGN, (r * s) % R>();
            });
            y[s] = acc;
        });
        static_for<0, R>([&](auto sc) { v[decltype(sc)::value] = y[decltype(sc)::value]; });
    }
};

template <typename T, int SIGN>
struct Dft<T, 2, SIGN> {
    __host__ __device__ __forceinline__ static void run(cpx<T> (&v)[2])
    {
        const cpx<T> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    }
};

template <typename T, int SIGN>
struct Dft<T, 3, SIGN> {
    __host__ __device__ __forceinline__ static void run(cpx<T> (&v)[3])
    {
        constexpr T h = static_cast<T>(0.5);
        constexpr T s3 = static_cast<T>(0.86602540378443864676);
        const cpx<T> t = v[1] + v[2];
        const cpx<T> u = mul_isign<T, SIGN>(v[1] - v[2], s3);
        const cpx<T> m = v[0] - scale(t, h);
        v[0] = v[0] + t;
        v[1] = m + u;
        v[2] = m - u;
    }
};

template <typename T, int SIGN>
struct Dft<T, 4, SIGN> {
    __host__ __device__ __forceinline__ static void run(cpx<T> (&v)[4])
    {
        const cpx<T> t0 = v[0] + v[2], t1 = v[0] - v[2];
        const cpx<T> t2 = v[1] + v[3];
        const cpx<T> t3 = mul_isign<T, SIGN>(v[1] - v[3], T(1));
        v[0] = t0 + t2;
        v[2] = t0 - t2;
        v[1] = t1 + t3;
        v[3] = t1 - t3;
    }
};

template <typename T, int SIGN>
struct Dft<T, 5, SIGN> {
    __host__ __device__ __forceinline__ static void run(cpx<T> (&v)[5])
    {
        constexpr ce_cs w1 = ce_cis(1, 5), w2 = ce_cis(2, 5);
        constexpr T c1 = static_cast<T>(w1.c), c2 = static_cast<T>(w2.c);
        constexpr T s1 = static_cast<T>(w1.s), s2 = static_cast<T>(w2.s);
        const cpx<T> a1 = v[1] + v[4], b1 = v[1] - v[4];
        const cpx<T> a2 = v[2] + v[3], b2 = v[2] - v[3];
        const cpx<T> m1 = v[0] + scale(a1, c1) + scale(a2, c2);
        const cpx<T> m2 = v[0] + scale(a1, c2) + scale(a2, c1);
        const cpx<T> n1 = mul_isign<T, SIGN>(scale(b1, s1) + scale(b2, s2), T(1));
        const cpx<T> n2 = mul_isign<T, SIGN>(scale(b1, s2) - scale(b2, s1), T(1));
        v[0] = v[0] + a1 + a2;
        v[1] = m1 + n1;
        v[4] = m1 - n1;
        v[2] = m2 + n2;
        v[3] = m2 - n2;
    }
};

template <typename T, int SIGN>
struct Dft<T, 8, SIGN> {
    __host__ __device__ __forceinline__ static void run(cpx<T> (&v)[8])
    {
        cpx<T> e[4] = {v[0], v[2], v[4], v[6]};
        cpx<T> o[4] = {v[1], v[3], v[5], v[7]};
        Dft<T, 4, SIGN>::run(e);
        Dft<T, 4, SIGN>::run(o);
        constexpr T r2 = static_cast<T>(0.70710678118654752440);
        // W_8^1 = (1 + SIGN i)/sqrt2, W_8^2 = SIGN i, W_8^3 = (-1 + SIGN i)/sqrt2
        const cpx<T> o1 = scale(o[1] + mul_isign<T, SIGN>(o[1], T(1)), r2);
        const cpx<T> o2 = mul_isign<T, SIGN>(o[2], T(1));
        const cpx<T> o3 = scale(mul_isign<T, SIGN>(o[3], T(1)) - o[3], r2);
        v[0] = e[0] + o[0];
        v[4] = e[0] - o[0];
        v[1] = e[1] + o1;
        v[5] = e[1] - o1;
        v[2] = e[2] + o2;
        v[6] = e[2] - o2;
        v[3] = e[3] + o3;
        v[7] = e[3] - o3;
    }
};

// ---------------------------------------------------------------------------
// Stockham autosort stages
// ---------------------------------------------------------------------------
constexpr int pick_radix(int m)
{
    if (m == 16) return 4;
    if (m % 8 == 0) return 8;
    if (m % 4 == 0) return 4;
    if (m % 2 == 0) return 2;
    if (m % 3 == 0) return 3;
    if (m % 5 == 0) return 5;
    return m;
}

// One stage: radix R, NS = product of the radices already applied.
template <typename T, int N, int R, int NS, int SIGN>
__host__ __device__ __forceinline__ void stockham_stage(const cpx<T> (&x)[N], cpx<T> (&y)[N])
{
    static_for<0, N / R>([&](auto jc) {
        constexpr int j = decltype(jc)::value;
        constexpr int k = j % NS;
        cpx<T> v[R];
        static_for<0, R>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            constexpr int m = (k * r * (N / (NS * R))) % N;
            if constexpr (m == 0)
                v[r] = x[j + r * (N / R)];
            else
                v[r] = x[j + r * (N / R)] * twiddle<T, N, SIGN, m>();
        });
        Dft<T, R, SIGN>::run(v);
        constexpr int idxD = (j / NS) * NS * R + k;
        static_for<0, R>([&](auto rc) {
            constexpr int r = decltype(rc)::value;
            y[idxD + r * NS] = v[r];
        });
    });
}

template <typename T, int N, int NS, int SIGN>
__host__ __device__ __forceinline__ void fft_stages(cpx<T> (&x)[N])
{
    if constexpr (NS < N) {
        constexpr int R = pick_radix(N / NS);
        cpx<T> y[N];
        stockham_stage<T, N, R, NS, SIGN>(x, y);
        fft_stages<T, N, NS * R, SIGN>(y);
        static_for<0, N>([&](auto ic) { x[decltype(ic)::value] = y[decltype(ic)::value]; });
    }
}

// ---------------------------------------------------------------------------
// Prime-factor (Good-Thomas) split N = N1 N2 with gcd(N1, N2) = 1: with the input
// index n = (N2 n1 + N1 n2) mod N and the output index k (k = k1 mod N1, k = k2 mod N2,
// Chinese remainder), exp(+-2 pi i n k / N) = W_N1^(n1 k1) W_N2^(n2 k2), so the length-N
// DFT is an N1 x N2 two-dimensional DFT with no twiddle factors between the stages.
// Both index maps are compile-time register renamings.  N1 is the power-of-two part.
// ---------------------------------------------------------------------------
constexpr int pow2_part(int n)
{
    int p = 1;
    while (n % 2 == 0) {
        n /= 2;
        p *= 2;
    }
    return p;
}
constexpr int crt_index(int k1, int k2, int n1, int n2)
{
    for (int k = 0; k < n1 * n2; ++k)
        if (k % n1 == k1 && k % n2 == k2) return k;
    return -1;
}

template <typename T, int N, int SIGN>
__host__ __device__ __forceinline__ void fft(cpx<T> (&x)[N]);

template <typename T, int N, int N1, int SIGN>
__host__ __device__ __forceinline__ void fft_pfa(cpx<T> (&x)[N])
{
    constexpr int N2 = N / N1;
    cpx<T> A[N];  // A[k1 * N2 + n2]
    static_for<0, N2>([&](auto n2c) {
        constexpr int n2 = decltype(n2c)::value;
        cpx<T> a[N1];
        static_for<0, N1>([&](auto n1c) {
            constexpr int n1 = decltype(n1c)::value;
            constexpr int n = (N2 * n1 + N1 * n2) % N;  // constant: keeps x in registers
            a[n1] = x[n];
        });
        fft<T, N1, SIGN>(a);
        static_for<0, N1>([&](auto k1c) {
            constexpr int k1 = decltype(k1c)::value;
            A[k1 * N2 + n2] = a[k1];
        });
    });
    static_for<0, N1>([&](auto k1c) {
        constexpr int k1 = decltype(k1c)::value;
        cpx<T> b[N2];
        static_for<0, N2>([&](auto n2c) {
            constexpr int n2 = decltype(n2c)::value;
            b[n2] = A[k1 * N2 + n2];
        });
        fft<T, N2, SIGN>(b);
        static_for<0, N2>([&](auto k2c) {
            constexpr int k2 = decltype(k2c)::value;
            constexpr int k = crt_index(k1, k2, N1, N2);
            x[k] = b[k2];
        });
    });
}

// In-place length-N DFT of a register array (unscaled, natural order in and out):
// prime-factor split when N has a power-of-two part and an odd part, else Stockham.
template <typename T, int N, int SIGN>
__host__ __device__ __forceinline__ void fft(cpx<T> (&x)[N])
{
    static_assert(SIGN == 1 || SIGN == -1, "sign");
    constexpr int N1 = pow2_part(N);
    if constexpr (N1 > 1 && N1 < N)
        fft_pfa<T, N, N1, SIGN>(x);
    else if constexpr (N > 1)
        fft_stages<T, N, 1, SIGN>(x);
}

}  // namespace tcx
