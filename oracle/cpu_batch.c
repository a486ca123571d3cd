/* cpu_batch.c -- CPU-baseline timing harness (test/bench infrastructure,
   never part of the product): runs a transform function of the reference
   minfft over a list of 1D lengths on T POSIX threads and returns the wall
   time.  Thread k takes entries k, k+T, k+2T, ... of the list, with its own
   aux per length (a reference aux owns scratch, minfft.c:17-22) and its own
   buffers, and passes over its entries `reps` times; the clock runs from
   the moment every thread has made its auxes to the moment the last one
   finishes.  The reference's functions are passed in as pointers (the
   caller loads oracle/_ref/libminfft_ref32_fast.so), so this file builds
   without the reference sources and calls them with no interpreter in the
   timed loop.  Complex data are interleaved pairs of `float` or `double`
   (elem_bytes 8 or 16 per complex point). */
#define _POSIX_C_SOURCE 200809L
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef void *(*mkaux_fn)(int d, int *Ns);
typedef void (*run_fn)(void *x, void *y, void *aux);
typedef void (*free_fn)(void *aux);

struct shared {
    int ready; /* threads whose auxes are made (atomic) */
    int go;    /* 1: run, 2: give up (atomic) */
};

struct job {
    const int *sizes;
    int count, threads, k, reps, elem_bytes;
    mkaux_fn mk;
    run_fn run;
    free_fn fr;
    struct shared *sh;
    int ok;
};

static int lg2(int n) {
    int l = 0;
    while (l < 31 && (1 << l) < n) ++l;
    return l;
}

static void pause_briefly(void) {
    struct timespec t = {0, 20000}; /* 20 us */
    nanosleep(&t, 0);
}

static void *worker(void *p) {
    struct job *j = (struct job *)p;
    void *aux[32];
    int i, r, maxn = 1, go;
    unsigned char *x, *y;
    unsigned s = 12345u + (unsigned)j->k;
    memset(aux, 0, sizeof(aux));
    for (i = j->k; i < j->count; i += j->threads)
        if (j->sizes[i] > maxn) maxn = j->sizes[i];
    x = (unsigned char *)malloc((size_t)maxn * (size_t)j->elem_bytes);
    y = (unsigned char *)malloc((size_t)maxn * (size_t)j->elem_bytes);
    j->ok = x && y;
    if (j->ok) {
        /* inputs in [-1, 1): a linear congruential sequence per thread */
        for (i = 0; i < maxn * 2; ++i) {
            s = s * 1664525u + 1013904223u;
            if (j->elem_bytes == 8)
                ((float *)x)[i] = (float)((double)(s >> 8) / 8388608.0 - 1.0);
            else
                ((double *)x)[i] = (double)(s >> 8) / 8388608.0 - 1.0;
        }
        for (i = j->k; i < j->count && j->ok; i += j->threads) {
            const int l = lg2(j->sizes[i]);
            if (j->sizes[i] < 1 || (1 << l) != j->sizes[i]) {
                j->ok = 0;
            } else if (!aux[l]) {
                int n = j->sizes[i];
                aux[l] = j->mk(1, &n);
                if (!aux[l]) j->ok = 0;
            }
        }
    }
    __atomic_add_fetch(&j->sh->ready, 1, __ATOMIC_SEQ_CST);
    while ((go = __atomic_load_n(&j->sh->go, __ATOMIC_SEQ_CST)) == 0) pause_briefly();
    if (go == 1 && j->ok)
        for (r = 0; r < j->reps; ++r)
            for (i = j->k; i < j->count; i += j->threads) j->run(x, y, aux[lg2(j->sizes[i])]);
    for (i = 0; i < 32; ++i)
        if (aux[i]) j->fr(aux[i]);
    free(x);
    free(y);
    return 0;
}

static double now_s(void) {
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)t.tv_sec + 1e-9 * (double)t.tv_nsec;
}

/* Wall seconds of `reps` passes over the entries on `threads` threads, or
   -1 on a failure (a thread that could not start, allocation, a maker
   returning NULL, a length that is not a power of two, bad arguments). */
double cpu_batch_seconds(const int *sizes, int count, int threads, int reps, int elem_bytes, mkaux_fn mk,
                         run_fn run, free_fn fr) {
    pthread_t tid[256];
    struct job jobs[256];
    struct shared sh;
    double t0, t1;
    int k, started = 0, ok = 1;
    if (!sizes || count < 1 || threads < 1 || threads > 256 || reps < 1 || !mk || !run || !fr ||
        (elem_bytes != 8 && elem_bytes != 16))
        return -1.0;
    sh.ready = 0;
    sh.go = 0;
    for (k = 0; k < threads; ++k) {
        jobs[k].sizes = sizes;
        jobs[k].count = count;
        jobs[k].threads = threads;
        jobs[k].k = k;
        jobs[k].reps = reps;
        jobs[k].elem_bytes = elem_bytes;
        jobs[k].mk = mk;
        jobs[k].run = run;
        jobs[k].fr = fr;
        jobs[k].sh = &sh;
        jobs[k].ok = 1;
        if (pthread_create(&tid[k], 0, worker, &jobs[k])) break;
        ++started;
    }
    /* every started thread has made its auxes (each counts itself once) */
    while (__atomic_load_n(&sh.ready, __ATOMIC_SEQ_CST) < started) pause_briefly();
    if (started < threads) ok = 0;
    for (k = 0; k < started; ++k)
        if (!jobs[k].ok) ok = 0;
    t0 = now_s();
    __atomic_store_n(&sh.go, ok ? 1 : 2, __ATOMIC_SEQ_CST);
    for (k = 0; k < started; ++k) pthread_join(tid[k], 0);
    t1 = now_s();
    return ok ? t1 - t0 : -1.0;
}
