/*
 * minfft_oracle.c -- CPU restatement of minfft's transform algorithms.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this file; only the test-suite, the smoke check and the CPU-baseline
 * leg of bench.py use it, and only as the checker.
 *
 * It restates, in plain C, the algorithms of the reference minfft.c
 * (aimukhin/minfft) so that the GPU path can be compared against the same
 * arithmetic on the same inputs:
 *   - twiddles      : octant-reduced cos/sin of 2*pi*n/N     (minfft.c:983-1018)
 *   - complex DFT   : split-radix 2/4 DIF recursion with
 *                     N=1,2,4,8 terminal codelets            (minfft.c:77-272, 293-488)
 *   - real DFT      : half-length complex DFT + post-twiddle (minfft.c:511-579)
 *   - inverse real  : pre-twiddle + half-length inverse DFT  (minfft.c:600-659)
 *   - DCT/DST-2     : reorder + real DFT + post-twiddle      (minfft.c:692-787)
 *   - DCT/DST-3     : pre-twiddle + inverse real + reorder   (minfft.c:790-887)
 *   - DCT/DST-4     : pre-twiddle + half-length DFT + post   (minfft.c:890-981)
 *   - d dimensions  : 1D routine along every axis, last axis
 *                     first, as mkcx/mkrx/realdft do         (minfft.c:33-72, 582-597, 662-687)
 *
 * Precision: `oreal` is double, or float when ORACLE_SINGLE is defined
 * (mirrors MINFFT_SINGLE, minfft.h:16-29).  Exported names carry an o64_ or
 * o32_ prefix so both builds can live in one process.
 *
 * Pinning: tests/test_oracle.py checks every function here against the
 * compiled reference (oracle/_ref), numpy/scipy FP64 transforms, closed
 * forms (delta, tone, constant) and long-double brute force for small N.
 */
#include <stdlib.h>
#include <string.h>
#include <math.h>

#ifdef ORACLE_SINGLE
typedef float oreal;
#define OPFX(name) o32_##name
#else
typedef double oreal;
#define OPFX(name) o64_##name
#endif

typedef struct { oreal re, im; } ocx;

static const oreal O_PI = 3.141592653589793238462643383279502884L;
static const oreal O_SQRT2 = 1.414213562373095048801688724209698079L;
static const oreal O_ISQRT2 = 0.707106781186547524400844362104849039L;

/* ------------------------------------------------------------------ */
/* roots of unity, minfft.c:983-1018: reduce n to [0, N/8] by the      */
/* symmetries of cos/sin, then call libm on 2*pi*n/N                   */
/* ------------------------------------------------------------------ */
static oreal rsin(long n, long N);
static oreal rcos(long n, long N)
{
	if (n < 0) return rcos(-n, N);
	if (n >= N / 2) return -rcos(n - N / 2, N);
	if (n >= N / 4) return -rsin(n - N / 4, N);
	if (n > N / 8) return rsin(N / 4 - n, N);
	return cos(2 * O_PI * n / N);
}
static oreal rsin(long n, long N)
{
	if (n < 0) return -rsin(-n, N);
	if (n >= N / 2) return -rsin(n - N / 2, N);
	if (n >= N / 4) return rcos(n - N / 4, N);
	if (n > N / 8) return rcos(N / 4 - n, N);
	return sin(2 * O_PI * n / N);
}
/* exp(2*pi*i*n/N) */
static ocx root(long n, long N)
{
	ocx w;
	w.re = rcos(n, N);
	w.im = rsin(n, N);
	return w;
}

static ocx cmul(ocx a, ocx b)
{
	ocx c;
	c.re = a.re * b.re - a.im * b.im;
	c.im = a.re * b.im + a.im * b.re;
	return c;
}

/* ------------------------------------------------------------------ */
/* split-radix complex DFT                                            */
/* ------------------------------------------------------------------ */

/* Per-level twiddle chain of minfft_mkaux_dft_1d (minfft.c:1075-1086):
   for each level M = N, N/2, ..., 16: for n < M/4 the pair
   (w^-n, w^-3n), w = exp(2*pi*i/M).  For the inverse transform the
   reference conjugates on use (minfft.c:473-482). */
static ocx *make_chain(long N)
{
	long M, n, cnt = 0;
	ocx *e, *p;
	for (M = N; M >= 16; M /= 2) cnt += M / 2;
	e = (ocx *)malloc((size_t)(cnt > 0 ? cnt : 1) * sizeof(ocx));
	if (!e) return NULL;
	p = e;
	for (M = N; M >= 16; M /= 2)
		for (n = 0; n < M / 4; ++n) {
			*p++ = root(-n, M);
			*p++ = root(-3 * n, M);
		}
	return e;
}

/* i*a (s=+1) or -i*a (s=-1) */
static ocx rot(ocx a, int s)
{
	ocx r;
	r.re = -s * a.im;
	r.im = s * a.re;
	return r;
}

/* Recursive split-radix DIF over N points: input x (contiguous), scratch t
   (N points), output y with stride sy.  sgn=-1 forward, +1 inverse.  The
   four leaves follow the reference's codelets (minfft.c:82-231): their
   operation order is kept so results agree to the last bit. */
static void sr_dft(long N, ocx *x, ocx *t, ocx *y, long sy, const ocx *e, int sgn)
{
	long n;
	if (N == 1) {
		y[0] = x[0];
		return;
	}
	if (N == 2) {
		ocx a = x[0], b = x[1], s, d;
		s.re = a.re + b.re; s.im = a.im + b.im;
		d.re = a.re - b.re; d.im = a.im - b.im;
		y[0] = s;
		y[sy] = d;
		return;
	}
	if (N == 4) {
		ocx p0, p1, q0, q1;
		/* p = even sums, q0 = x0-x2, q1 = (sgn i)*... rotated odd difference */
		p0.re = x[0].re + x[2].re; p0.im = x[0].im + x[2].im;
		p1.re = x[1].re + x[3].re; p1.im = x[1].im + x[3].im;
		q0.re = x[0].re - x[2].re; q0.im = x[0].im - x[2].im;
		{
			ocx dd;
			dd.re = x[1].re - x[3].re; dd.im = x[1].im - x[3].im;
			/* forward: t3 = I*(x1-x3); inverse uses -I (conjugate codelet) */
			q1.re = -dd.im; q1.im = dd.re;
			if (sgn > 0) { q1.re = dd.im; q1.im = -dd.re; }
		}
		y[0].re = p0.re + p1.re;       y[0].im = p0.im + p1.im;
		y[sy].re = q0.re - q1.re;      y[sy].im = q0.im - q1.im;
		y[2 * sy].re = p0.re - p1.re;  y[2 * sy].im = p0.im - p1.im;
		y[3 * sy].re = q0.re + q1.re;  y[3 * sy].im = q0.im + q1.im;
		return;
	}
	if (N == 8) {
		/* two radix-4 halves (even / odd inputs), then combine with
		   the 8th roots of unity written with 1/sqrt(2) */
		ocx a[4], b[4], tt, u, v;
		int h;
		for (h = 0; h < 2; ++h) {
			ocx s0, s1, d0, d1, dd, *o = h ? b : a;
			s0.re = x[h].re + x[h + 4].re;  s0.im = x[h].im + x[h + 4].im;
			s1.re = x[h + 2].re + x[h + 6].re; s1.im = x[h + 2].im + x[h + 6].im;
			d0.re = x[h].re - x[h + 4].re;  d0.im = x[h].im - x[h + 4].im;
			dd.re = x[h + 2].re - x[h + 6].re; dd.im = x[h + 2].im - x[h + 6].im;
			if (sgn < 0) { d1.re = -dd.im; d1.im = dd.re; }
			else { d1.re = dd.im; d1.im = -dd.re; }
			o[0].re = s0.re + s1.re; o[0].im = s0.im + s1.im;
			o[1].re = d0.re - d1.re; o[1].im = d0.im - d1.im;
			o[2].re = s0.re - s1.re; o[2].im = s0.im - s1.im;
			o[3].re = d0.re + d1.re; o[3].im = d0.im + d1.im;
		}
		/* b1 *= w8^(sgn), b2 *= w8^(2 sgn), b3 *= w8^(3 sgn) */
		tt = b[1];
		if (sgn < 0) {
			u.re = O_ISQRT2 * (tt.re + tt.im);
			u.im = O_ISQRT2 * (tt.im - tt.re);
		} else {
			u.re = O_ISQRT2 * (tt.re - tt.im);
			u.im = O_ISQRT2 * (tt.im + tt.re);
		}
		b[1] = u;
		tt = b[2];
		if (sgn < 0) { v.re = tt.im; v.im = -tt.re; }
		else { v.re = -tt.im; v.im = tt.re; }
		b[2] = v;
		tt = b[3];
		if (sgn < 0) {
			u.re = O_ISQRT2 * (tt.im - tt.re);
			u.im = -O_ISQRT2 * (tt.im + tt.re);
		} else {
			u.re = -O_ISQRT2 * (tt.re + tt.im);
			u.im = O_ISQRT2 * (tt.re - tt.im);
		}
		b[3] = u;
		for (h = 0; h < 4; ++h) {
			y[h * sy].re = a[h].re + b[h].re;
			y[h * sy].im = a[h].im + b[h].im;
			y[(h + 4) * sy].re = a[h].re - b[h].re;
			y[(h + 4) * sy].im = a[h].im - b[h].im;
		}
		return;
	}
	/* DIF step (minfft.c:236-267): t[n]=x[n]+x[n+N/2], t[n+N/4]=x[n+N/4]+x[n+3N/4],
	   t[n+N/2]=(d0 - s*i*d1)*e1, t[n+3N/4]=(d0 + s*i*d1)*e3 */
	for (n = 0; n < N / 4; ++n) {
		ocx a0 = x[n], a1 = x[n + N / 4], a2 = x[n + N / 2], a3 = x[n + 3 * N / 4];
		ocx s0, s1, d0, d1, lo, hi, w1, w3;
		s0.re = a0.re + a2.re; s0.im = a0.im + a2.im;
		s1.re = a1.re + a3.re; s1.im = a1.im + a3.im;
		d0.re = a0.re - a2.re; d0.im = a0.im - a2.im;
		{
			ocx dd;
			dd.re = a1.re - a3.re; dd.im = a1.im - a3.im;
			d1 = rot(dd, sgn < 0 ? 1 : -1);
		}
		t[n] = s0;
		t[n + N / 4] = s1;
		lo.re = d0.re - d1.re; lo.im = d0.im - d1.im;
		hi.re = d0.re + d1.re; hi.im = d0.im + d1.im;
		w1 = e[2 * n];
		w3 = e[2 * n + 1];
		if (sgn > 0) { w1.im = -w1.im; w3.im = -w3.im; }
		t[n + N / 2] = cmul(lo, w1);
		t[n + 3 * N / 4] = cmul(hi, w3);
	}
	sr_dft(N / 2, t, t, y, 2 * sy, e + N / 2, sgn);
	sr_dft(N / 4, t + N / 2, t + N / 2, y + sy, 4 * sy, e + 3 * N / 4, sgn);
	sr_dft(N / 4, t + 3 * N / 4, t + 3 * N / 4, y + 3 * sy, 4 * sy, e + 3 * N / 4, sgn);
}

/* A 1D complex transform of length N along a strided line. */
typedef struct {
	long N;
	ocx *e;   /* split-radix chain */
	ocx *t;   /* scratch, N points */
	ocx *in;  /* contiguous copy of the line */
} cplan;

static int cplan_init(cplan *p, long N)
{
	p->N = N;
	p->e = make_chain(N);
	p->t = (ocx *)malloc((size_t)N * sizeof(ocx));
	p->in = (ocx *)malloc((size_t)N * sizeof(ocx));
	return p->e && p->t && p->in;
}
static void cplan_free(cplan *p)
{
	free(p->e);
	free(p->t);
	free(p->in);
}
/* out[k*so] = sum_n in[n*si] exp(sgn*2*pi*i*n*k/N) */
static void cplan_run(cplan *p, const ocx *x, long si, ocx *y, long so, int sgn)
{
	long n;
	for (n = 0; n < p->N; ++n) p->in[n] = x[n * si];
	sr_dft(p->N, p->in, p->t, y, so, p->e, sgn);
}

/* ------------------------------------------------------------------ */
/* d-dimensional helpers: apply a 1D line transform along one axis     */
/* ------------------------------------------------------------------ */
static long prod(int d, const int *Ns)
{
	long p = 1;
	int i;
	for (i = 0; i < d; ++i) p *= Ns[i];
	return p;
}

static int valid_dims(int d, const int *Ns)
{
	int i;
	if (d < 1) return 0;
	for (i = 0; i < d; ++i)
		if (Ns[i] <= 0 || (Ns[i] & (Ns[i] - 1))) return 0;
	return 1;
}

/* complex transform along axis ax of the array a (in place through a
   line buffer) whose dims are Ns[0..d-1] */
static int cx_axis(int d, const int *Ns, int ax, ocx *a, int sgn)
{
	long L = Ns[ax], stride = 1, outer, o, q;
	int i;
	cplan p;
	ocx *line;
	for (i = ax + 1; i < d; ++i) stride *= Ns[i];
	outer = prod(d, Ns) / (L * stride);
	if (!cplan_init(&p, L)) { cplan_free(&p); return -1; }
	line = (ocx *)malloc((size_t)L * sizeof(ocx));
	if (!line) { cplan_free(&p); return -1; }
	for (o = 0; o < outer; ++o)
		for (q = 0; q < stride; ++q) {
			ocx *base = a + o * L * stride + q;
			long k;
			cplan_run(&p, base, stride, line, 1, sgn);
			for (k = 0; k < L; ++k) base[k * stride] = line[k];
		}
	free(line);
	cplan_free(&p);
	return 0;
}

/* ------------------------------------------------------------------ */
/* exported: complex DFT (minfft_dft / minfft_invdft, minfft.c:287-290, */
/* 503-506); unnormalised; axes transformed last to first like mkcx    */
/* ------------------------------------------------------------------ */
int OPFX(dft)(int d, const int *Ns, const oreal *x, oreal *y, int inverse)
{
	long P;
	int ax;
	if (!valid_dims(d, Ns)) return -1;
	P = prod(d, Ns);
	if ((const void *)x != (void *)y) memmove(y, x, (size_t)P * sizeof(ocx));
	for (ax = d - 1; ax >= 0; --ax)
		if (cx_axis(d, Ns, ax, (ocx *)y, inverse ? 1 : -1)) return -1;
	return 0;
}

/* ------------------------------------------------------------------ */
/* real DFT, 1D (minfft.c:511-579)                                     */
/* ------------------------------------------------------------------ */
static int realdft_line(long N, const oreal *x, ocx *z, long sz)
{
	long n, L = N / 2;
	cplan p;
	ocx *w, *t;
	if (N == 1) {
		z[0].re = x[0];
		z[0].im = 0;
		return 0;
	}
	if (N == 2) {
		z[0].re = x[0] + x[1];
		z[0].im = 0;
		z[sz].re = x[0] - x[1];
		z[sz].im = 0;
		return 0;
	}
	if (!cplan_init(&p, L)) { cplan_free(&p); return -1; }
	w = (ocx *)malloc((size_t)L * sizeof(ocx) * 2);
	if (!w) { cplan_free(&p); return -1; }
	t = w + L;
	for (n = 0; n < L; ++n) {
		w[n].re = x[2 * n];
		w[n].im = x[2 * n + 1];
	}
	cplan_run(&p, w, 1, t, 1, -1);
	/* DC and Nyquist */
	z[0].re = t[0].re + t[0].im;
	z[0].im = 0;
	z[sz * L].re = t[0].re - t[0].im;
	z[sz * L].im = 0;
	for (n = 1; n < L / 2; ++n) {
		/* u = (t[n] + conj t[L-n])/2 ; v = (t[n] - conj t[L-n]) * e / (2i),
		   e = exp(-2 pi i n / N) */
		ocx a = t[n], b = t[L - n], u, v, dlt, w1;
		ocx em = root(-n, N);
		u.re = (a.re + b.re) / 2;
		u.im = (a.im - b.im) / 2;
		dlt.re = a.re - b.re;
		dlt.im = a.im + b.im;
		w1.re = em.im;  /* e / i */
		w1.im = -em.re;
		v.re = (dlt.re * w1.re - dlt.im * w1.im) / 2;
		v.im = (dlt.re * w1.im + dlt.im * w1.re) / 2;
		z[sz * n].re = u.re + v.re;
		z[sz * n].im = u.im + v.im;
		z[sz * (L - n)].re = u.re - v.re;
		z[sz * (L - n)].im = -u.im + v.im;
	}
	z[sz * (L / 2)].re = t[L / 2].re;
	z[sz * (L / 2)].im = -t[L / 2].im;
	free(w);
	cplan_free(&p);
	return 0;
}

/* inverse real DFT, 1D (minfft.c:600-659), unnormalised */
static int invrealdft_line(long N, const ocx *z, oreal *y)
{
	long n, L = N / 2;
	cplan p;
	ocx *t, *w;
	if (N == 1) {
		y[0] = z[0].re;
		return 0;
	}
	if (N == 2) {
		y[0] = z[0].re + z[1].re;
		y[1] = z[0].re - z[1].re;
		return 0;
	}
	if (!cplan_init(&p, L)) { cplan_free(&p); return -1; }
	t = (ocx *)malloc((size_t)L * sizeof(ocx) * 2);
	if (!t) { cplan_free(&p); return -1; }
	w = t + L;
	/* imaginary parts of z[0] and z[L] are ignored (minfft.c:629-630) */
	t[0].re = z[0].re + z[L].re;
	t[0].im = z[0].re - z[L].re;
	for (n = 1; n < L / 2; ++n) {
		/* u = z[n] + conj z[L-n] ; v = i*(z[n] - conj z[L-n])*conj(e),
		   e = exp(-2 pi i n / N) */
		ocx a = z[n], b = z[L - n], u, v, dlt, te;
		ocx em = root(-n, N);
		u.re = a.re + b.re;
		u.im = a.im - b.im;
		dlt.re = a.re - b.re;
		dlt.im = a.im + b.im;
		te.re = em.im;  /* i*conj(e) */
		te.im = em.re;
		v.re = dlt.re * te.re - dlt.im * te.im;
		v.im = dlt.re * te.im + dlt.im * te.re;
		t[n].re = u.re + v.re;
		t[n].im = u.im + v.im;
		t[L - n].re = u.re - v.re;
		t[L - n].im = -u.im + v.im;
	}
	t[L / 2].re = 2 * z[L / 2].re;
	t[L / 2].im = -2 * z[L / 2].im;
	cplan_run(&p, t, 1, w, 1, 1);
	for (n = 0; n < L; ++n) {
		y[2 * n] = w[n].re;
		y[2 * n + 1] = w[n].im;
	}
	free(t);
	cplan_free(&p);
	return 0;
}

/* ------------------------------------------------------------------ */
/* exported: real DFT, d dimensions (minfft.c:582-597): real lines     */
/* along the last axis, then complex DFT over the leading axes of the  */
/* N1 x ... x (Nd/2+1) result                                          */
/* ------------------------------------------------------------------ */
int OPFX(realdft)(int d, const int *Ns, const oreal *x, oreal *z)
{
	long Nd, H, rows, r;
	int ax, *Hs;
	ocx *zc = (ocx *)z;
	oreal *xc;
	if (!valid_dims(d, Ns)) return -1;
	Nd = Ns[d - 1];
	H = Nd / 2 + 1;
	rows = prod(d - 1, Ns);
	/* copy input first so that in-place calls behave */
	xc = (oreal *)malloc((size_t)(rows * Nd) * sizeof(oreal));
	if (!xc) return -1;
	memcpy(xc, x, (size_t)(rows * Nd) * sizeof(oreal));
	for (r = 0; r < rows; ++r)
		if (realdft_line(Nd, xc + r * Nd, zc + r * H, 1)) { free(xc); return -1; }
	free(xc);
	if (d == 1) return 0;
	Hs = (int *)malloc((size_t)d * sizeof(int));
	if (!Hs) return -1;
	memcpy(Hs, Ns, (size_t)d * sizeof(int));
	Hs[d - 1] = (int)H;
	for (ax = d - 2; ax >= 0; --ax)
		if (cx_axis(d, Hs, ax, zc, -1)) { free(Hs); return -1; }
	free(Hs);
	return 0;
}

/* exported: inverse real DFT, d dimensions (minfft.c:662-687).  Works on
   a copy, so z is preserved (the reference clobbers it for d > 1). */
int OPFX(invrealdft)(int d, const int *Ns, const oreal *z, oreal *y)
{
	long Nd, H, rows, r;
	int ax, *Hs;
	ocx *t;
	if (!valid_dims(d, Ns)) return -1;
	Nd = Ns[d - 1];
	H = Nd / 2 + 1;
	rows = prod(d - 1, Ns);
	t = (ocx *)malloc((size_t)(rows * H) * sizeof(ocx));
	if (!t) return -1;
	memcpy(t, z, (size_t)(rows * H) * sizeof(ocx));
	if (d > 1) {
		Hs = (int *)malloc((size_t)d * sizeof(int));
		if (!Hs) { free(t); return -1; }
		memcpy(Hs, Ns, (size_t)d * sizeof(int));
		Hs[d - 1] = (int)H;
		for (ax = d - 2; ax >= 0; --ax)
			if (cx_axis(d, Hs, ax, t, 1)) { free(Hs); free(t); return -1; }
		free(Hs);
	}
	for (r = 0; r < rows; ++r)
		if (invrealdft_line(Nd, t + r * H, y + r * Nd)) { free(t); return -1; }
	free(t);
	return 0;
}

/* ------------------------------------------------------------------ */
/* real symmetric transforms, 1D lines                                */
/* kind: 2 = DCT-2, 3 = DCT-3, 4 = DCT-4, -2/-3/-4 = DST-2/3/4         */
/* ------------------------------------------------------------------ */
static int r2r_line(int kind, long N, const oreal *x, oreal *y)
{
	long n, L = N / 2;
	int sgn = kind < 0 ? -1 : 1, typ = kind < 0 ? -kind : kind;
	if (N == 1) {
		/* minfft.c:700-704, 749-753, 798-802, 848-852, 897-901, 944-948 */
		if (typ == 2) y[0] = 2 * x[0];
		else if (typ == 3) y[0] = x[0];
		else y[0] = O_SQRT2 * x[0];
		return 0;
	}
	if (typ == 2) {
		/* minfft.c:707-725 (DCT-2), 756-774 (DST-2) */
		oreal *t = (oreal *)malloc((size_t)(N + 2) * sizeof(oreal));
		ocx *z = (ocx *)t;
		if (!t) return -1;
		for (n = 0; n < L; ++n) {
			t[n] = x[2 * n];
			t[L + n] = sgn * x[N - 1 - 2 * n];
		}
		{
			oreal *tc = (oreal *)malloc((size_t)N * sizeof(oreal));
			if (!tc) { free(t); return -1; }
			memcpy(tc, t, (size_t)N * sizeof(oreal));
			if (realdft_line(N, tc, z, 1)) { free(tc); free(t); return -1; }
			free(tc);
		}
		if (sgn > 0) {
			y[0] = 2 * z[0].re;
			for (n = 1; n < L; ++n) {
				ocx e = root(-n, 4 * N);
				y[n] = 2 * (z[n].re * e.re - z[n].im * e.im);
				y[N - n] = -2 * (z[n].re * e.im + z[n].im * e.re);
			}
			y[L] = O_SQRT2 * z[L].re;
		} else {
			y[N - 1] = 2 * z[0].re;
			for (n = 1; n < L; ++n) {
				ocx e = root(-n, 4 * N);
				y[n - 1] = -2 * (z[n].re * e.im + z[n].im * e.re);
				y[N - n - 1] = 2 * (z[n].re * e.re - z[n].im * e.im);
			}
			y[L - 1] = O_SQRT2 * z[L].re;
		}
		free(t);
		return 0;
	}
	if (typ == 3) {
		/* minfft.c:807-824 (DCT-3), 857-874 (DST-3) */
		ocx *z = (ocx *)malloc((size_t)(L + 1) * sizeof(ocx));
		oreal *t = (oreal *)malloc((size_t)N * sizeof(oreal));
		if (!z || !t) { free(z); free(t); return -1; }
		if (sgn > 0) {
			z[0].re = x[0];
			z[0].im = 0;
			for (n = 1; n < L; ++n) {
				ocx e = root(-n, 4 * N);
				z[n].re = x[n] * e.re - x[N - n] * e.im;
				z[n].im = -x[n] * e.im - x[N - n] * e.re;
			}
			z[L].re = O_SQRT2 * x[L];
			z[L].im = 0;
		} else {
			z[0].re = x[N - 1];
			z[0].im = 0;
			for (n = 1; n < L; ++n) {
				ocx e = root(-n, 4 * N);
				z[n].re = x[N - n - 1] * e.re - x[n - 1] * e.im;
				z[n].im = -x[N - n - 1] * e.im - x[n - 1] * e.re;
			}
			z[L].re = O_SQRT2 * x[L - 1];
			z[L].im = 0;
		}
		if (invrealdft_line(N, z, t)) { free(z); free(t); return -1; }
		for (n = 0; n < L; ++n) {
			y[2 * n] = t[n];
			y[N - 1 - 2 * n] = sgn * t[L + n];
		}
		free(z);
		free(t);
		return 0;
	}
	/* typ == 4: minfft.c:906-921 (DCT-4), 953-968 (DST-4) */
	{
		cplan p;
		ocx *t = (ocx *)malloc((size_t)L * sizeof(ocx) * 2), *T;
		if (!t) return -1;
		T = t + L;
		if (!cplan_init(&p, L)) { cplan_free(&p); free(t); return -1; }
		for (n = 0; n < L; ++n) {
			ocx e = root(-n, 2 * N);
			if (sgn > 0) {
				t[n].re = e.re * x[2 * n] - e.im * x[N - 1 - 2 * n];
				t[n].im = e.re * x[N - 1 - 2 * n] + e.im * x[2 * n];
			} else {
				t[n].re = -e.re * x[2 * n] - e.im * x[N - 1 - 2 * n];
				t[n].im = e.re * x[N - 1 - 2 * n] - e.im * x[2 * n];
			}
		}
		cplan_run(&p, t, 1, T, 1, -1);
		for (n = 0; n < L; ++n) {
			ocx e0 = root(-(4 * n + 1), 8 * N), e1 = root(-(4 * n + 3), 8 * N);
			if (sgn > 0) {
				y[2 * n] = 2 * (e0.re * T[n].re - e0.im * T[n].im);
				y[2 * n + 1] = 2 * (e1.re * T[L - 1 - n].re + e1.im * T[L - 1 - n].im);
			} else {
				y[2 * n] = 2 * (e0.re * T[n].im + e0.im * T[n].re);
				y[2 * n + 1] = 2 * (-e1.re * T[L - 1 - n].im + e1.im * T[L - 1 - n].re);
			}
		}
		cplan_free(&p);
		free(t);
		return 0;
	}
}

/* exported: DCT/DST of types 2,3,4 in d dimensions (mkrx, minfft.c:57-72):
   the 1D transform along every axis, last axis first */
int OPFX(r2r)(int kind, int d, const int *Ns, const oreal *x, oreal *y)
{
	long P;
	int ax, typ = kind < 0 ? -kind : kind;
	oreal *line_in, *line_out;
	if (!valid_dims(d, Ns) || typ < 2 || typ > 4) return -1;
	P = prod(d, Ns);
	if ((const void *)x != (void *)y) memmove(y, x, (size_t)P * sizeof(oreal));
	for (ax = d - 1; ax >= 0; --ax) {
		long L = Ns[ax], stride = 1, outer, o, q, k;
		int i;
		for (i = ax + 1; i < d; ++i) stride *= Ns[i];
		outer = P / (L * stride);
		line_in = (oreal *)malloc((size_t)L * sizeof(oreal) * 2);
		if (!line_in) return -1;
		line_out = line_in + L;
		for (o = 0; o < outer; ++o)
			for (q = 0; q < stride; ++q) {
				oreal *base = y + o * L * stride + q;
				for (k = 0; k < L; ++k) line_in[k] = base[k * stride];
				if (r2r_line(kind, L, line_in, line_out)) { free(line_in); return -1; }
				for (k = 0; k < L; ++k) base[k * stride] = line_out[k];
			}
		free(line_in);
	}
	return 0;
}

/* ------------------------------------------------------------------ */
/* brute-force O(N^2) complex DFT in long double (checker of checkers) */
/* ------------------------------------------------------------------ */
int OPFX(brute_dft)(int N, const oreal *x, oreal *y, int inverse)
{
	int k, n;
	const long double tp = 6.283185307179586476925286766559005768L;
	for (k = 0; k < N; ++k) {
		long double sr = 0, si = 0;
		for (n = 0; n < N; ++n) {
			long m = ((long)n * k) % N;
			long double ang = (inverse ? tp : -tp) * (long double)m / N;
			long double c = cosl(ang), s = sinl(ang);
			long double ar = x[2 * n], ai = x[2 * n + 1];
			sr += ar * c - ai * s;
			si += ar * s + ai * c;
		}
		y[2 * k] = (oreal)sr;
		y[2 * k + 1] = (oreal)si;
	}
	return 0;
}
