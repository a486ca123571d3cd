// host_checks.cpp -- drives the host code of the library (plans, batch
// plans, staging of host arrays, schedules, the tuning cache, error and
// allocation-failure paths, streams) so that a build of that code under
// AddressSanitizer / UndefinedBehaviorSanitizer sees every path run.  The
// kernels are compiled as usual; only host code is instrumented.
//
// Built twice, like the client programs: -DMINFFT_SINGLE=0 against
// libminfft.so, -DMINFFT_SINGLE=1 against libminfft_single.so.  Prints
// "ok" last when every check held.  With no GPU visible it checks that the
// makers fail with a reason and that nothing else runs.
#include <hip/hip_runtime_api.h>
#include <sanitizer/lsan_interface.h>
#include <unistd.h>

#include <atomic>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "minfft.h"
#include "minfft_ext.h"

typedef std::complex<minfft_real> cx;

static std::atomic<int> g_fail(0);
static const double kPi = 3.14159265358979323846;

#define CHECK(c, ...)                                  \
    do {                                               \
        if (!(c)) {                                    \
            std::printf("FAIL %s:%d: ", __FILE__, __LINE__); \
            std::printf(__VA_ARGS__);                  \
            std::printf("\n");                         \
            ++g_fail;                                  \
        }                                              \
    } while (0)

static const double kTol = sizeof(minfft_real) == 8 ? 1e-11 : 2e-4;

static double rel_err(const minfft_real* a, const minfft_real* b, size_t n, double scale) {
    double num = 0, den = 0;
    for (size_t i = 0; i < n; ++i) {
        const double d = (double)a[i] - scale * (double)b[i];
        num += d * d;
        den += scale * scale * (double)b[i] * (double)b[i];
    }
    return den > 0 ? std::sqrt(num / den) : std::sqrt(num);
}

static thread_local unsigned g_seed = 12345u;
static minfft_real urand() {
    g_seed = g_seed * 1664525u + 1013904223u;
    return (minfft_real)((g_seed >> 8) * (1.0 / 16777216.0) - 0.5);
}

// d-dimensional makers by family, through the general maker
static minfft_aux* make(int fam, int d, int* n) {
    switch (fam) {
        case 0: return minfft_mkaux_dft(d, n);
        case 1: return minfft_mkaux_realdft(d, n);
        case 2: return minfft_mkaux_t2t3(d, n);
        default: return minfft_mkaux_t4(d, n);
    }
}

// one round trip of each kind of family `fam` on host arrays of dims n:
// invdft(dft(x)) = P x, invrealdft(realdft(x)) = P x, dct3(dct2(x)) = 2^d P x
// (and dst), dct4(dct4(x)) = 2^d P x (and dst); in place and out of place
static void round_trips(int fam, int d, int* n) {
    long long P = 1;
    for (int i = 0; i < d; ++i) P *= n[i];
    minfft_aux* a = make(fam, d, n);
    CHECK(a, "maker fam %d d %d: %s", fam, d, minfft_last_error());
    if (!a) return;
    const size_t nr = (size_t)P;
    const size_t nh = (size_t)(P / n[d - 1] * (n[d - 1] / 2 + 1));
    std::vector<minfft_real> x(2 * nr), y(2 * nr), z(2 * nr);
    for (auto& v : x) v = urand();
    if (fam == 0) {
        minfft_dft((minfft_cmpl*)x.data(), (minfft_cmpl*)y.data(), a);
        minfft_invdft((minfft_cmpl*)y.data(), (minfft_cmpl*)z.data(), a);
        CHECK(rel_err(z.data(), x.data(), 2 * nr, (double)P) < kTol, "dft round trip d=%d P=%lld", d, P);
        z = x;  // in place
        minfft_dft((minfft_cmpl*)z.data(), (minfft_cmpl*)z.data(), a);
        CHECK(rel_err(z.data(), y.data(), 2 * nr, 1.0) < kTol, "dft in place d=%d P=%lld", d, P);
    } else if (fam == 1) {
        minfft_realdft(x.data(), (minfft_cmpl*)y.data(), a);
        minfft_invrealdft((minfft_cmpl*)y.data(), z.data(), a);
        CHECK(rel_err(z.data(), x.data(), nr, (double)P) < kTol, "realdft round trip d=%d P=%lld", d, P);
        (void)nh;
    } else {
        const double s = std::ldexp((double)P, d);
        void (*fw[2])(minfft_real*, minfft_real*, const minfft_aux*);
        void (*bw[2])(minfft_real*, minfft_real*, const minfft_aux*);
        if (fam == 2) {
            fw[0] = minfft_dct2, bw[0] = minfft_dct3, fw[1] = minfft_dst2, bw[1] = minfft_dst3;
        } else {
            fw[0] = bw[0] = minfft_dct4, fw[1] = bw[1] = minfft_dst4;
        }
        for (int k = 0; k < 2; ++k) {
            fw[k](x.data(), y.data(), a);
            bw[k](y.data(), z.data(), a);
            CHECK(rel_err(z.data(), x.data(), nr, s) < kTol, "fam %d kind %d round trip d=%d P=%lld", fam, k, d, P);
        }
    }
    minfft_free_aux(a);
}

static void error_paths() {
    int bad[3] = {4, 3, 8};
    CHECK(!minfft_mkaux_dft_1d(0) && *minfft_last_error(), "N=0 accepted");
    CHECK(!minfft_mkaux_dft_1d(-8), "N=-8 accepted");
    CHECK(!minfft_mkaux_realdft_1d(12), "N=12 accepted");
    CHECK(!minfft_mkaux_t2t3(3, bad), "3 in dims accepted");
    CHECK(!minfft_mkaux_t4(0, bad), "d=0 accepted");
    CHECK(!minfft_mkaux_dft(1, NULL), "NULL dims accepted");
    minfft_free_aux(NULL);
    minfft_batch_entry e[2];
    std::memset(e, 0, sizeof(e));
    e[0].kind = MINFFT_DFT, e[0].d = 1, e[0].n[0] = 16;
    e[1] = e[0];
    CHECK(!minfft_mkaux_batch(0, e, MINFFT_FP32), "empty batch accepted");
    CHECK(!minfft_mkaux_batch(2, e, 7), "precision 7 accepted");
    e[1].kind = 10;
    CHECK(!minfft_mkaux_batch(2, e, MINFFT_FP32) && *minfft_last_error(), "kind 10 accepted");
    e[1].kind = MINFFT_DCT2, e[1].d = 4;
    CHECK(!minfft_mkaux_batch(2, e, MINFFT_FP32), "d=4 accepted");
    e[1].d = 2, e[1].n[1] = 6;
    CHECK(!minfft_mkaux_batch(2, e, MINFFT_FP32), "length 6 accepted");
    e[1].n[1] = 8, e[1].in_offset = 2;
    CHECK(!minfft_mkaux_batch(2, e, MINFFT_FP32), "misaligned offset accepted");
    CHECK(minfft_enqueue(MINFFT_DFT, NULL, NULL, NULL, NULL) == -1, "NULL aux enqueued");
    CHECK(minfft_set_stream(NULL, NULL) == -1, "NULL aux took a stream");
    char buf[8];
    CHECK(minfft_batch_describe(NULL, buf, (int)sizeof(buf)) == -1, "NULL plan described");
    CHECK(minfft_batch_schedule(NULL, 0, "") == -1, "NULL plan scheduled");
}

// host arrays of several sizes: the zero-copy path (<= 4096 points), one
// bounce piece, the pipelined bounce (> 4 MiB), pageable copies (> 64 MiB);
// long lengths through the two-step and pre/post kernels
static void host_sizes() {
    const int lens[] = {1, 2, 8, 4096, 1 << 15, 1 << 18, 1 << 20, 1 << 22, 1 << 23};
    for (int N : lens) {
        int n[1] = {N};
        round_trips(0, 1, n);
        if (N <= (1 << 22)) round_trips(1, 1, n);
        if (N <= (1 << 20)) round_trips(2, 1, n), round_trips(3, 1, n);
    }
    int n2[2] = {64, 32}, n3[3] = {8, 4, 16}, big[2] = {2048, 2048}, one[3] = {1, 1, 2};
    for (int f = 0; f < 4; ++f) {
        round_trips(f, 2, n2);
        round_trips(f, 3, n3);
        round_trips(f, 3, one);
    }
    round_trips(1, 2, big);
    round_trips(2, 2, big);
}

// the 1D DFT against a direct sum (N = 64)
static void direct_sum() {
    const int N = 64;
    minfft_aux* a = minfft_mkaux_dft_1d(N);
    CHECK(a, "mkaux_dft_1d(64): %s", minfft_last_error());
    if (!a) return;
    std::vector<cx> x(N), y(N);
    for (auto& v : x) v = cx(urand(), urand());
    minfft_dft((minfft_cmpl*)x.data(), (minfft_cmpl*)y.data(), a);
    double err = 0, ref = 0;
    for (int k = 0; k < N; ++k) {
        std::complex<double> s = 0;
        for (int m = 0; m < N; ++m)
            s += std::complex<double>(x[m]) * std::polar(1.0, -2 * kPi * (double)((long long)k * m % N) / N);
        err += std::norm(s - std::complex<double>(y[k]));
        ref += std::norm(s);
    }
    CHECK(std::sqrt(err / ref) < kTol, "dft vs direct sum: %g", std::sqrt(err / ref));
    minfft_free_aux(a);
}

// device arrays: synchronous calls, minfft_enqueue on a caller's stream,
// minfft_set_stream, minfft_enqueue_noalloc
static void device_paths() {
    const int N = 1 << 16;
    minfft_aux* a = minfft_mkaux_dft_1d(N);
    CHECK(a, "mkaux_dft_1d: %s", minfft_last_error());
    if (!a) return;
    const size_t bytes = (size_t)N * sizeof(cx);
    std::vector<cx> x(N), y(N), z(N);
    for (auto& v : x) v = cx(urand(), urand());
    minfft_dft((minfft_cmpl*)x.data(), (minfft_cmpl*)y.data(), a);  // reference: host arrays
    void *dx = nullptr, *dy = nullptr;
    hipStream_t s = nullptr;
    if (hipMalloc(&dx, bytes) != hipSuccess || hipMalloc(&dy, bytes) != hipSuccess ||
        hipStreamCreate(&s) != hipSuccess) {
        CHECK(false, "hip allocation");
        return;
    }
    (void)hipMemcpy(dx, x.data(), bytes, hipMemcpyHostToDevice);
    auto fetch = [&](const char* what) {
        (void)hipMemcpy(z.data(), dy, bytes, hipMemcpyDeviceToHost);
        CHECK(rel_err((minfft_real*)z.data(), (minfft_real*)y.data(), 2 * (size_t)N, 1.0) < kTol, "%s", what);
        (void)hipMemset(dy, 0, bytes);
    };
    minfft_dft((minfft_cmpl*)dx, (minfft_cmpl*)dy, a);
    fetch("synchronous call on device arrays");
    CHECK(minfft_enqueue(MINFFT_DFT, dx, dy, a, s) == 0, "enqueue: %s", minfft_last_error());
    (void)hipStreamSynchronize(s);
    fetch("enqueue on a caller's stream");
    CHECK(minfft_enqueue(MINFFT_DFT, x.data(), dy, a, s) == -1, "host pointer enqueued");
    CHECK(minfft_enqueue(42, dx, dy, a, s) == -1, "kind 42 enqueued");
    CHECK(minfft_set_stream(a, s) == 0, "set_stream: %s", minfft_last_error());
    minfft_dft((minfft_cmpl*)dx, (minfft_cmpl*)dy, a);
    fetch("synchronous call on a set stream");
    minfft_dft((minfft_cmpl*)x.data(), (minfft_cmpl*)z.data(), a);
    CHECK(rel_err((minfft_real*)z.data(), (minfft_real*)y.data(), 2 * (size_t)N, 1.0) < kTol,
          "host arrays on a set stream");
    CHECK(minfft_enqueue_noalloc(MINFFT_DFT, dx, dy, a, NULL) == 0, "noalloc: %s", minfft_last_error());
    (void)hipStreamSynchronize(s);
    fetch("enqueue_noalloc");
    CHECK(minfft_set_stream(a, NULL) == 0, "set_stream(NULL)");
    minfft_aux* huge = minfft_mkaux_dft_1d(1 << 24);
    if (huge) {
        CHECK(minfft_enqueue_noalloc(MINFFT_DFT, dx, dy, huge, NULL) == -1, "noalloc over its scratch");
        minfft_free_aux(huge);
    }
    (void)hipFree(dx);
    (void)hipFree(dy);
    (void)hipStreamDestroy(s);
    minfft_free_aux(a);
}

// batch plans: groups of every launch form, host and device arrays,
// describe, schedule, the tuning cache file
static void batch_paths(const char* cache) {
    std::vector<minfft_batch_entry> e;
    long long off = 0;
    const int prec = sizeof(minfft_real) == 8 ? MINFFT_FP64 : MINFFT_FP32;
    const size_t isz = sizeof(minfft_real);
    auto add = [&](int kind, int d, int n0, int n1, int n2) {
        minfft_batch_entry b;
        std::memset(&b, 0, sizeof(b));
        b.kind = kind, b.d = d, b.n[0] = n0, b.n[1] = n1, b.n[2] = n2;
        long long P = n0 * (long long)(d > 1 ? n1 : 1) * (d > 2 ? n2 : 1);
        b.in_offset = off, b.out_offset = off;
        off += 2 * P * (long long)isz + 256;  // room for a complex or half-spectrum output
        e.push_back(b);
    };
    for (int i = 0; i < 24; ++i) add(i % 2 ? MINFFT_INVDFT : MINFFT_DFT, 1, 16 << (i % 8), 0, 0);
    for (int i = 0; i < 6; ++i) add(MINFFT_DFT, 1, 1 << 16, 0, 0);
    for (int i = 0; i < 4; ++i) add(MINFFT_DFT, 1, 1 << 15, 0, 0);
    add(MINFFT_REALDFT, 2, 64, 64, 0);
    add(MINFFT_DCT2, 2, 128, 32, 0);
    add(MINFFT_DST4, 3, 8, 8, 8);
    add(MINFFT_INVREALDFT, 1, 1 << 18, 0, 0);
    add(MINFFT_DCT3, 1, 1 << 17, 0, 0);
    setenv("MINFFT_TUNE_CACHE", cache, 1);
    minfft_batch_plan* p = minfft_mkaux_batch((int)e.size(), e.data(), prec);
    CHECK(p, "mkaux_batch: %s", minfft_last_error());
    if (!p) return;
    const size_t bytes = (size_t)off;
    std::vector<minfft_real> x(bytes / isz), y(bytes / isz), z(bytes / isz);
    for (auto& v : x) v = urand();
    CHECK(minfft_batch(p, x.data(), bytes, y.data(), bytes) == 0, "batch host: %s", minfft_last_error());
    // an input buffer ending inside the last entry's input
    CHECK(minfft_batch(p, x.data(), (size_t)e.back().in_offset + 8, y.data(), bytes) == -1,
          "short input accepted");
    CHECK(minfft_batch(p, x.data(), bytes, y.data(), (size_t)e.back().out_offset + 8) == -1,
          "short output accepted");
    char small[16];
    const int need = minfft_batch_describe(p, small, (int)sizeof(small));
    CHECK(need > (int)sizeof(small) && std::strlen(small) == sizeof(small) - 1, "describe truncation");
    std::string full((size_t)need, '\0');
    CHECK(minfft_batch_describe(p, &full[0], need) == need, "describe");
    CHECK(minfft_batch_schedule(p, 0, "tpw=2,bogus=1") == -1, "unknown key accepted");
    CHECK(minfft_batch_schedule(p, 9999, "") == -1, "group 9999 accepted");
    CHECK(minfft_batch_schedule(p, 1, "e32=1") == 0, "schedule: %s", minfft_last_error());
    void *dx = nullptr, *dy = nullptr;
    if (hipMalloc(&dx, bytes) == hipSuccess && hipMalloc(&dy, bytes) == hipSuccess) {
        (void)hipMemcpy(dx, x.data(), bytes, hipMemcpyHostToDevice);
        (void)hipMemcpy(dy, y.data(), bytes, hipMemcpyHostToDevice);
        for (int r = 0; r < 3; ++r) CHECK(minfft_batch_run(p, dx, dy, NULL) == 0, "batch_run: %s", minfft_last_error());
        (void)hipDeviceSynchronize();
        (void)hipMemcpy(z.data(), dy, bytes, hipMemcpyDeviceToHost);
        CHECK(rel_err(z.data(), y.data(), z.size(), 1.0) < kTol, "batch_run vs batch");
        CHECK(minfft_batch_run(p, x.data(), dy, NULL) == -1, "host pointer run");
    }
    (void)hipFree(dx);
    (void)hipFree(dy);
    minfft_free_batch(p);
    // a second plan of the same shapes reads the cache file written above
    p = minfft_mkaux_batch((int)e.size(), e.data(), prec);
    CHECK(p, "mkaux_batch again: %s", minfft_last_error());
    if (p) {
        CHECK(minfft_batch(p, x.data(), bytes, z.data(), bytes) == 0, "batch again");
        CHECK(rel_err(z.data(), y.data(), z.size(), 1.0) < kTol, "cached schedule result");
        minfft_free_batch(p);
    }
    unsetenv("MINFFT_TUNE_CACHE");
}

// every allocation of a maker and of a first call failing in turn: the
// maker returns NULL or the call reports an error; nothing leaks
static void allocation_failures() {
    char v[16];
    for (int k = 1; k <= 12; ++k) {
        std::snprintf(v, sizeof(v), "%d", k);
        setenv("MINFFT_FAIL_ALLOC", v, 1);
        int n[2] = {256, 512};
        minfft_aux* a = minfft_mkaux_realdft(2, n);
        if (a) {
            std::vector<minfft_real> x(256 * 512, 1), y(256 * 514);
            minfft_realdft(x.data(), (minfft_cmpl*)y.data(), a);
            minfft_free_aux(a);
        }
        minfft_batch_entry e[3];
        std::memset(e, 0, sizeof(e));
        e[0].kind = MINFFT_DFT, e[0].d = 1, e[0].n[0] = 1 << 16;
        e[1] = e[0], e[1].in_offset = e[1].out_offset = (1 << 16) * 2 * (long long)sizeof(minfft_real);
        e[2].kind = MINFFT_DCT2, e[2].d = 1, e[2].n[0] = 1 << 18;
        // the same k written differently restarts the count for the batch maker
        std::snprintf(v, sizeof(v), "0%d", k);
        setenv("MINFFT_FAIL_ALLOC", v, 1);
        minfft_batch_plan* p = minfft_mkaux_batch(3, e, sizeof(minfft_real) == 8 ? MINFFT_FP64 : MINFFT_FP32);
        if (p) minfft_free_batch(p);
        unsetenv("MINFFT_FAIL_ALLOC");
    }
}

// distinct aux objects on two threads at once
static void threads() {
    std::thread t[2];
    for (int i = 0; i < 2; ++i)
        t[i] = std::thread([i] {
            int n[2] = {128, 64 << i};
            for (int r = 0; r < 3; ++r) round_trips(i ? 2 : 0, 2, n);
        });
    for (auto& th : t) th.join();
}

// Leak check now, then leave without running the HIP runtime's exit-time
// teardown: the sanitizer runtime's allocator checks fail inside it, which
// says nothing about this library.
static int finish(int rc) {
    std::fflush(stdout);
    __lsan_do_leak_check();
    std::fflush(stdout);
    std::fflush(stderr);
    _exit(rc);
}

int main(int argc, char** argv) {
    setvbuf(stdout, nullptr, _IOLBF, 0);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        CHECK(!minfft_mkaux_dft_1d(16), "maker succeeded without a device");
        CHECK(std::strstr(minfft_last_error(), "device") != NULL, "reason: %s", minfft_last_error());
        minfft_batch_entry e;
        std::memset(&e, 0, sizeof(e));
        e.kind = MINFFT_DFT, e.d = 1, e.n[0] = 8;
        CHECK(!minfft_mkaux_batch(1, &e, MINFFT_FP32), "batch maker succeeded without a device");
        error_paths();
        std::printf(g_fail ? "FAILED (%d)\n" : "no device: ok\n", g_fail.load());
        return finish(g_fail != 0);
    }
    const char* cache = argc > 1 ? argv[1] : "host_checks.tune";
    const struct {
        const char* name;
        void (*run)(const char*);
    } sections[] = {
        {"error paths", [](const char*) { error_paths(); }},
        {"direct sum", [](const char*) { direct_sum(); }},
        {"host array sizes", [](const char*) { host_sizes(); }},
        {"device arrays and streams", [](const char*) { device_paths(); }},
        {"batch plans and tuning cache", [](const char* c) { batch_paths(c); }},
        {"allocation failures", [](const char*) { allocation_failures(); }},
        {"threads", [](const char*) { threads(); }},
    };
    for (const auto& sec : sections) {
        std::printf("[%s]\n", sec.name);
        sec.run(cache);
    }
    if (g_fail) {
        std::printf("FAILED (%d)\n", g_fail.load());
        return finish(1);
    }
    std::printf("ok\n");
    return finish(0);
}
