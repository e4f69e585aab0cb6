// Host check of the register FFT (tc_amd/csrc/fft_reg.hpp) that every GPU pass runs:
// each length and sign against a direct DFT in long double on seeded inputs, in FP64
// (relative error <= 1e-14) and FP32 (<= 2e-6).  The code under test is the same
// __host__ __device__ template the kernels instantiate; the prime-factor split, the
// Stockham stages and the split-pencil sub-transforms all pass through fft<T, N, SIGN>.
// Exit status 0 when every length passes.
#include <cmath>
#include <cstdio>
#include <random>

#include "fft_reg.hpp"

using namespace tcx;

template <typename T, int N, int SIGN>
double rel_error(unsigned seed)
{
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd;
    double worst = 0.0;
    for (int trial = 0; trial < 4; ++trial) {
        cpx<T> x[N];
        long double xr[N], xi[N];
        for (int i = 0; i < N; ++i) {
            x[i] = cpx<T>{(T)nd(rng), (T)nd(rng)};
            xr[i] = x[i].re;
            xi[i] = x[i].im;
        }
        fft<T, N, SIGN>(x);
        double err = 0.0, nrm = 0.0;
        for (int k = 0; k < N; ++k) {
            long double sr = 0.0L, si = 0.0L;
            for (int n = 0; n < N; ++n) {
                const long double a =
                    SIGN * 2.0L * 3.141592653589793238462643383279502884L * (long double)((long)n * k % N) / N;
                sr += xr[n] * cosl(a) - xi[n] * sinl(a);
                si += xr[n] * sinl(a) + xi[n] * cosl(a);
            }
            err = std::fmax(err, (double)hypotl((long double)x[k].re - sr, (long double)x[k].im - si));
            nrm = std::fmax(nrm, (double)hypotl(sr, si));
        }
        worst = std::fmax(worst, err / nrm);
    }
    return worst;
}

int failures = 0;

template <int N>
void check()
{
    const double d0 = rel_error<double, N, -1>(17 * N + 1), d1 = rel_error<double, N, 1>(17 * N + 2);
    const double f0 = rel_error<float, N, -1>(17 * N + 3), f1 = rel_error<float, N, 1>(17 * N + 4);
    const bool ok = d0 <= 1e-14 && d1 <= 1e-14 && f0 <= 2e-6 && f1 <= 2e-6;
    failures += !ok;
    std::printf("N=%2d fp64 %.2e %.2e fp32 %.2e %.2e %s\n", N, d0, d1, f0, f1, ok ? "ok" : "FAIL");
}

int main()
{
    // grid sides and the sub-lengths of the split pencils (NZ/2, NZ/4)
    check<2>(); check<3>(); check<4>(); check<5>(); check<6>(); check<8>(); check<9>(); check<10>();
    check<12>(); check<16>(); check<18>(); check<20>(); check<24>(); check<32>(); check<36>(); check<40>();
    check<48>();
    return failures == 0 ? 0 : 1;
}
