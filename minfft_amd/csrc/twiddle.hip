// twiddle.hip -- device generation of the twiddle table in FP64.
//
// tw[M + m] = exp(-2*pi*i*m/M) for every power of two M <= Mmax and m < M
// (entry 0 unused), i.e. 2*Mmax points.  The angle is reduced to the first
// octant with exact integer arithmetic on m -- the symmetries the reference
// uses in ncos/nsin (minfft.c:983-1018) -- and evaluated with sincospi on the
// exact dyadic argument 2*m/M, so every entry is the correctly reduced value
// (cos/sin of multiples of pi/4 come out exact).  The FP32 table is the FP64
// one rounded to nearest.
#include <hip/hip_runtime.h>

namespace mfx {

// cos and sin of 2*pi*n/M for 0 <= n < M, via octant reduction
__device__ static void root_cs(long long n, long long M, double& c, double& s) {
    // exp(+2*pi*i*n/M); reduce n into [0, M/8] tracking the symmetry
    // quadrant: n = q*M/4 + r
    const long long Q = M / 4;
    double cc, ss;
    if (M < 8) {
        // M in {1,2,4}: angles are multiples of 90 degrees
        long long k = (n * 4) / M;  // quarter turns
        const double tc[4] = {1.0, 0.0, -1.0, 0.0};
        const double ts[4] = {0.0, 1.0, 0.0, -1.0};
        c = tc[k & 3];
        s = ts[k & 3];
        return;
    }
    long long q = n / Q, r = n - q * Q;
    // angle inside the quadrant: r in [0, M/4)
    bool swap = false;
    if (r > M / 8) { r = Q - r; swap = true; }
    sincospi(2.0 * (double)r / (double)M, &ss, &cc);
    if (r == 0) { ss = 0.0; cc = 1.0; }
    double c0 = swap ? ss : cc, s0 = swap ? cc : ss;  // cos/sin of the in-quadrant angle
    switch (q & 3) {
        case 0: c = c0; s = s0; break;
        case 1: c = -s0; s = c0; break;
        case 2: c = -c0; s = -s0; break;
        default: c = s0; s = -c0; break;
    }
}

__global__ void twiddle_kernel(double2* __restrict__ tw64, float2* __restrict__ tw32, long long total) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    if (i == 0) {
        tw64[0] = make_double2(1.0, 0.0);
        if (tw32) tw32[0] = make_float2(1.0f, 0.0f);
        return;
    }
    // M = highest power of two <= i
    int lg = 63 - __clzll((unsigned long long)i);
    long long M = 1LL << lg;
    long long m = i - M;
    double c, s;
    root_cs(m, M, c, s);
    // exp(-2*pi*i*m/M) = cos - i sin
    double2 w = make_double2(c, -s);
    tw64[i] = w;
    if (tw32) tw32[i] = make_float2((float)w.x, (float)w.y);
}

// Completion word for the host: written once every earlier launch of the
// stream has finished (stream order), after a system-scope fence, into
// coherent pinned host memory, so a host thread can poll it instead of
// synchronising the stream.
__global__ void host_flag_kernel(unsigned* flag, unsigned v) {
    __threadfence_system();
    __atomic_store_n(flag, v, __ATOMIC_RELEASE);
}

hipError_t launch_host_flag(unsigned* flag, unsigned v, hipStream_t st) {
    (void)hipGetLastError();
    hipLaunchKernelGGL(host_flag_kernel, dim3(1), dim3(1), 0, st, flag, v);
    return hipGetLastError();
}

hipError_t launch_twiddles(double2* tw64, float2* tw32, long long total, hipStream_t st) {
    const int bs = 256;
    long long grid = (total + bs - 1) / bs;
    (void)hipGetLastError();  // a failed call before this one is not this launch's error
    hipLaunchKernelGGL(twiddle_kernel, dim3((unsigned)grid), dim3(bs), 0, st, tw64, tw32, total);
    return hipGetLastError();
}

}  // namespace mfx
