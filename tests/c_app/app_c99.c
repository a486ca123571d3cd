/* A C99 program that uses minfft.h with C99 complex numbers: <complex.h>
 * included first, so minfft_cmpl is `double complex` (`float complex` with
 * MINFFT_SINGLE), as the reference header allows (minfft.h:16-29).  Checks
 * a 2D DFT against a direct double sum and the real-DFT round trip; prints
 * "ok" and exits 0 on success. */
#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "minfft.h"

#if MINFFT_SINGLE
#define TOL 2e-5
#else
#define TOL 1e-12
#endif

int main(void) {
    enum { N1 = 8, N2 = 16, N = N1 * N2 };
    const double pi = 3.14159265358979323846;
    minfft_cmpl *x = malloc(N * sizeof *x), *y = malloc(N * sizeof *y);
    minfft_real *r = malloc(N * sizeof *r), *rb = malloc(N * sizeof *rb);
    minfft_cmpl *z = malloc(N1 * (N2 / 2 + 1) * sizeof *z);
    double num = 0, den = 0;
    int i, j, k, l;
    minfft_aux *a = minfft_mkaux_dft_2d(N1, N2), *ar = minfft_mkaux_realdft_2d(N1, N2);
    if (!x || !y || !r || !rb || !z || !a || !ar) return 2;
    for (i = 0; i < N; ++i) {
        x[i] = (minfft_real)(sin(0.37 * i) + 0.25) + (minfft_real)cos(1.3 * i) * I;
        r[i] = (minfft_real)sin(0.11 * i * i);
    }
    minfft_dft(x, y, a);
    for (k = 0; k < N1; ++k)
        for (l = 0; l < N2; ++l) {
            double complex s = 0;
            for (i = 0; i < N1; ++i)
                for (j = 0; j < N2; ++j)
                    s += x[i * N2 + j] * cexp(-2 * pi * I * ((double)i * k / N1 + (double)j * l / N2));
            num += pow(cabs(y[k * N2 + l] - s), 2);
            den += pow(cabs(s), 2);
        }
    if (sqrt(num / den) > TOL) { printf("dft error %g\n", sqrt(num / den)); return 1; }
    minfft_realdft(r, z, ar);
    minfft_invrealdft(z, rb, ar);
    num = den = 0;
    for (i = 0; i < N; ++i) {
        num += pow(rb[i] / N - r[i], 2);
        den += pow(r[i], 2);
    }
    if (sqrt(num / den) > TOL) { printf("realdft round trip error %g\n", sqrt(num / den)); return 1; }
    minfft_free_aux(a);
    minfft_free_aux(ar);
    free(x); free(y); free(r); free(rb); free(z);
    printf("ok\n");
    return 0;
}
