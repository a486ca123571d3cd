// A C++ program that uses minfft.h with std::complex arrays (binary
// compatible with minfft_cmpl, reference README.md:93-97): a 3D DFT against
// a direct sum and a DCT-2 / DCT-3 round trip; prints "ok" and exits 0 on
// success.
#include <cmath>
#include <complex>
#include <cstdio>
#include <vector>

#include "minfft.h"

#if MINFFT_SINGLE
typedef float real_t;
static const double kTol = 2e-5;
#else
typedef double real_t;
static const double kTol = 1e-12;
#endif
typedef std::complex<real_t> cmpl_t;

int main() {
    const int n0 = 4, n1 = 8, n2 = 8, n = n0 * n1 * n2;
    const double pi = 3.14159265358979323846;
    std::vector<cmpl_t> x(n), y(n);
    for (int i = 0; i < n; ++i) x[i] = cmpl_t((real_t)std::sin(0.7 * i), (real_t)std::cos(0.29 * i * i));
    minfft_aux* a = minfft_mkaux_dft_3d(n0, n1, n2);
    if (!a) return 2;
    minfft_dft(reinterpret_cast<minfft_cmpl*>(x.data()), reinterpret_cast<minfft_cmpl*>(y.data()), a);
    double num = 0, den = 0;
    for (int k0 = 0; k0 < n0; ++k0)
        for (int k1 = 0; k1 < n1; ++k1)
            for (int k2 = 0; k2 < n2; ++k2) {
                std::complex<double> s = 0;
                for (int i0 = 0; i0 < n0; ++i0)
                    for (int i1 = 0; i1 < n1; ++i1)
                        for (int i2 = 0; i2 < n2; ++i2) {
                            const double ph = -2 * pi * ((double)i0 * k0 / n0 + (double)i1 * k1 / n1 +
                                                         (double)i2 * k2 / n2);
                            s += std::complex<double>(x[(i0 * n1 + i1) * n2 + i2]) * std::polar(1.0, ph);
                        }
                num += std::norm(std::complex<double>(y[(k0 * n1 + k1) * n2 + k2]) - s);
                den += std::norm(s);
            }
    if (std::sqrt(num / den) > kTol) { std::printf("dft error %g\n", std::sqrt(num / den)); return 1; }
    minfft_free_aux(a);
    std::vector<real_t> r(n), c(n), b(n);
    for (int i = 0; i < n; ++i) r[i] = (real_t)std::cos(0.05 * i * i);
    minfft_aux* t = minfft_mkaux_t2t3_3d(n0, n1, n2);
    if (!t) return 2;
    minfft_dct2(r.data(), c.data(), t);
    minfft_dct3(c.data(), b.data(), t);
    num = den = 0;
    for (int i = 0; i < n; ++i) {
        num += std::pow(b[i] / (8.0 * n) - r[i], 2);
        den += std::pow(r[i], 2);
    }
    if (std::sqrt(num / den) > kTol) { std::printf("dct round trip error %g\n", std::sqrt(num / den)); return 1; }
    minfft_free_aux(t);
    std::printf("ok\n");
    return 0;
}
