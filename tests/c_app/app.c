/* app.c -- a program written against the reference's minfft.h API, as a
   user of the reference would write it (C89, host arrays), built unchanged
   against the MI355X libraries (libminfft.so, or libminfft_single.so with
   -DMINFFT_SINGLE=1).  Checks the transforms by their defining identities:
   the complex DFT against a direct O(N^2) sum, and round trips
   (inverse(forward(x)) = scale * x) of the complex DFT in 1D and 3D (in
   place), the real DFT in 2D, DCT-2/DCT-3, DST-2/DST-3 and DCT-4/DST-4;
   and that a maker rejects a length that is not a power of two.  Prints
   the errors and exits 0 when every one is within the bar. */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "minfft.h"

#if MINFFT_SINGLE
#define TOL 2e-5
#else
#define TOL 1e-12
#endif

static unsigned long state = 12345UL;

static double uni(void) {
    state = state * 6364136223846793005UL + 1442695040888963407UL;
    return (double)((state >> 11) & 0xFFFFFFFFFFFFFUL) / 4503599627370496.0 * 2.0 - 1.0;
}

static int fails = 0;

static void report(const char *what, double err) {
    printf("%-28s %.3e\n", what, err);
    if (!(err <= TOL)) fails++;
}

/* relative max error of a - s*b over n reals */
static double relerr(const minfft_real *a, const minfft_real *b, int n, double s) {
    double e = 0, m = 0;
    int i;
    for (i = 0; i < n; i++) {
        double d = fabs((double)a[i] - s * (double)b[i]);
        double v = fabs(s * (double)b[i]);
        if (d > e) e = d;
        if (v > m) m = v;
    }
    return m > 0 ? e / m : e;
}

static void complex_dft_direct(void) {
    enum { N = 64 };
    static minfft_cmpl x[N], y[N];
    static minfft_real ref[2 * N];
    const double pi = 3.14159265358979323846;
    minfft_aux *a = minfft_mkaux_dft_1d(N);
    int n, k;
    if (!a) { printf("minfft_mkaux_dft_1d failed\n"); fails++; return; }
    for (n = 0; n < N; n++) { x[n][0] = (minfft_real)uni(); x[n][1] = (minfft_real)uni(); }
    for (k = 0; k < N; k++) {
        double re = 0, im = 0;
        for (n = 0; n < N; n++) {
            double t = -2 * pi * (double)((n * k) % N) / N;
            re += x[n][0] * cos(t) - x[n][1] * sin(t);
            im += x[n][0] * sin(t) + x[n][1] * cos(t);
        }
        ref[2 * k] = (minfft_real)re;
        ref[2 * k + 1] = (minfft_real)im;
    }
    minfft_dft(x, y, a);
    report("dft 64 vs direct sum", relerr(&y[0][0], ref, 2 * N, 1.0));
    minfft_free_aux(a);
}

static void complex_round_trips(void) {
    enum { N = 1024, D0 = 8, D1 = 4, D2 = 16 };
    static minfft_cmpl x[N], y[N], z[N], w[D0 * D1 * D2], w0[D0 * D1 * D2];
    minfft_aux *a = minfft_mkaux_dft_1d(N), *b = minfft_mkaux_dft_3d(D0, D1, D2);
    int n;
    if (!a || !b) { printf("complex makers failed\n"); fails++; return; }
    for (n = 0; n < N; n++) { x[n][0] = (minfft_real)uni(); x[n][1] = (minfft_real)uni(); }
    minfft_dft(x, y, a);
    minfft_invdft(y, z, a);
    report("dft/invdft 1024", relerr(&z[0][0], &x[0][0], 2 * N, N));
    for (n = 0; n < D0 * D1 * D2; n++) {
        w[n][0] = w0[n][0] = (minfft_real)uni();
        w[n][1] = w0[n][1] = (minfft_real)uni();
    }
    minfft_dft(w, w, b); /* in place */
    minfft_invdft(w, w, b);
    report("dft/invdft 8x4x16 in place", relerr(&w[0][0], &w0[0][0], 2 * D0 * D1 * D2, D0 * D1 * D2));
    minfft_free_aux(a);
    minfft_free_aux(b);
}

static void real_round_trips(void) {
    enum { R = 32, C = 16, N = 128 };
    static minfft_real x[R * C], y[R * C], u[N], v[N], t[N];
    static minfft_cmpl z[R * (C / 2 + 1)];
    minfft_aux *a = minfft_mkaux_realdft_2d(R, C), *b = minfft_mkaux_t2t3_1d(N), *c = minfft_mkaux_t4_1d(N);
    int n;
    if (!a || !b || !c) { printf("real makers failed\n"); fails++; return; }
    for (n = 0; n < R * C; n++) x[n] = (minfft_real)uni();
    minfft_realdft(x, z, a);
    minfft_invrealdft(z, y, a);
    report("realdft/invrealdft 32x16", relerr(y, x, R * C, R * C));
    for (n = 0; n < N; n++) u[n] = (minfft_real)uni();
    minfft_dct2(u, v, b);
    minfft_dct3(v, t, b);
    report("dct2/dct3 128", relerr(t, u, N, 2 * N));
    minfft_dst2(u, v, b);
    minfft_dst3(v, t, b);
    report("dst2/dst3 128", relerr(t, u, N, 2 * N));
    minfft_dct4(u, v, c);
    minfft_dct4(v, t, c);
    report("dct4/dct4 128", relerr(t, u, N, 2 * N));
    minfft_dst4(u, v, c);
    minfft_dst4(v, t, c);
    report("dst4/dst4 128", relerr(t, u, N, 2 * N));
    minfft_free_aux(a);
    minfft_free_aux(b);
    minfft_free_aux(c);
}

int main(void) {
    complex_dft_direct();
    complex_round_trips();
    real_round_trips();
    if (minfft_mkaux_dft_1d(12) != NULL) {
        printf("minfft_mkaux_dft_1d(12) should fail\n");
        fails++;
    }
    minfft_free_aux(NULL); /* a no-op, as in the reference */
    printf(fails ? "FAILED %d\n" : "ok\n", fails);
    return fails ? 1 : 0;
}
