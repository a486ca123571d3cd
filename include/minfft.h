/*
 * minfft.h -- C interface of the MI355X (gfx950) minfft drop-in.
 *
 * Same 27 entry points, argument meaning and NULL-on-error convention as the
 * reference header (aimukhin/minfft minfft.h:35-63); a program built against
 * the reference links unchanged against libminfft.so (double precision) or,
 * with MINFFT_SINGLE defined, libminfft_single.so (single precision).
 *
 * Types (reference minfft.h:14-31): minfft_real is double, or float when
 * MINFFT_SINGLE is nonzero; minfft_cmpl is C99 complex when `complex` is
 * defined before inclusion, else an array of two minfft_real -- both layouts
 * are two adjacent reals.  minfft_aux is opaque; here it owns device
 * resources (twiddle table, scratch, a HIP stream).
 *
 * Pointers: x and y may be host memory (staged through the device) or device
 * memory from hipMalloc / managed memory (used in place).  Every call returns
 * with the result written.  Transforms are unnormalised, dims are in C order
 * (last index fastest) and all lengths are powers of two.
 */
#ifndef MINFFT_H
#define MINFFT_H

#ifdef __cplusplus
extern "C" {
#endif

#ifdef complex
#if MINFFT_SINGLE
typedef float minfft_real;
typedef float complex minfft_cmpl;
#else
typedef double minfft_real;
typedef double complex minfft_cmpl;
#endif
#else
#if MINFFT_SINGLE
typedef float minfft_real;
#else
typedef double minfft_real;
#endif
typedef minfft_real minfft_cmpl[2];
#endif

typedef struct minfft_aux minfft_aux;

/* transforms (reference minfft.h:35-44) */
/* forward complex DFT, y_k = sum_n x_n exp(-2 pi i n k / N)   -- minfft.c:287-290 */
void minfft_dft(minfft_cmpl *x, minfft_cmpl *y, const minfft_aux *a);
/* inverse complex DFT, unnormalised                          -- minfft.c:503-506 */
void minfft_invdft(minfft_cmpl *x, minfft_cmpl *y, const minfft_aux *a);
/* real DFT, N1 x ... x Nd reals -> N1 x ... x (Nd/2+1)       -- minfft.c:582-597 */
void minfft_realdft(minfft_real *x, minfft_cmpl *z, const minfft_aux *a);
/* inverse real DFT, unnormalised; z is preserved here        -- minfft.c:662-687 */
void minfft_invrealdft(minfft_cmpl *z, minfft_real *y, const minfft_aux *a);
/* DCT-2 / DST-2                                              -- minfft.c:735-738, 784-787 */
void minfft_dct2(minfft_real *x, minfft_real *y, const minfft_aux *a);
void minfft_dst2(minfft_real *x, minfft_real *y, const minfft_aux *a);
/* DCT-3 / DST-3                                              -- minfft.c:834-837, 884-887 */
void minfft_dct3(minfft_real *x, minfft_real *y, const minfft_aux *a);
void minfft_dst3(minfft_real *x, minfft_real *y, const minfft_aux *a);
/* DCT-4 / DST-4                                              -- minfft.c:931-934, 978-981 */
void minfft_dct4(minfft_real *x, minfft_real *y, const minfft_aux *a);
void minfft_dst4(minfft_real *x, minfft_real *y, const minfft_aux *a);

/* aux makers (reference minfft.h:46-61); NULL on invalid lengths, d < 1,
   allocation failure or no GPU (reason: minfft_last_error(), minfft_ext.h) */
minfft_aux *minfft_mkaux_dft_1d(int N);                       /* minfft.c:1057 */
minfft_aux *minfft_mkaux_dft_2d(int N1, int N2);              /* minfft.c:1104 */
minfft_aux *minfft_mkaux_dft_3d(int N1, int N2, int N3);      /* minfft.c:1112 */
minfft_aux *minfft_mkaux_dft(int d, int *Ns);                 /* minfft.c:1098 */
minfft_aux *minfft_mkaux_realdft_1d(int N);                   /* minfft.c:1121 */
minfft_aux *minfft_mkaux_realdft_2d(int N1, int N2);          /* minfft.c:1192 */
minfft_aux *minfft_mkaux_realdft_3d(int N1, int N2, int N3);  /* minfft.c:1200 */
minfft_aux *minfft_mkaux_realdft(int d, int *Ns);             /* minfft.c:1159 */
minfft_aux *minfft_mkaux_t2t3_1d(int N);                      /* minfft.c:1209 */
minfft_aux *minfft_mkaux_t2t3_2d(int N1, int N2);             /* minfft.c:1254 */
minfft_aux *minfft_mkaux_t2t3_3d(int N1, int N2, int N3);     /* minfft.c:1262 */
minfft_aux *minfft_mkaux_t2t3(int d, int *Ns);                /* minfft.c:1248 */
minfft_aux *minfft_mkaux_t4_1d(int N);                        /* minfft.c:1271 */
minfft_aux *minfft_mkaux_t4_2d(int N1, int N2);               /* minfft.c:1321 */
minfft_aux *minfft_mkaux_t4_3d(int N1, int N2, int N3);       /* minfft.c:1329 */
minfft_aux *minfft_mkaux_t4(int d, int *Ns);                  /* minfft.c:1315 */

/* release an aux; NULL is a no-op (reference minfft.h:63, minfft.c:1338-1347) */
void minfft_free_aux(minfft_aux *a);

#ifdef __cplusplus
}
#endif

#endif /* MINFFT_H */
