// noalloc.cpp -- minfft_enqueue_noalloc (minfft_ext.h): minfft_enqueue whose
// kernels take their scratch from a fixed device array of this library
// instead of the aux's own buffers, and which allocates nothing, for callers
// that cannot allocate device memory during the call (code under HIP graph
// capture, tight serving loops).
//
// Scratch a transform's kernels use: the two-step intermediate of axes
// longer than one work-group holds (FP32 > 2^14, FP64 > 2^13 points), the
// complex staging of real/DCT/DST axes longer than that, and the full-array
// temp of the multi-dimensional inverse real DFT (and of the real DFT with
// overlapping input and output).  All of it is written before it is read,
// so nothing needs clearing between calls.  Every call on one GPU shares
// that GPU's array: calls that overlap in time on one GPU (two host threads,
// or two streams whose work runs at once) race on it.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "engine.hpp"
#include "minfft_ext.h"

// the largest scratch a call may use (each GPU the library runs on holds
// one such array until the process ends)
static constexpr size_t kScratchBytes = MINFFT_NOALLOC_SCRATCH_BYTES;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
void* g_addr[kMaxDevices];  // device address of `scratch`, per device

// Address of this device's copy of `scratch`, looked up once per device
// (with that device current); nullptr with the error set on failure.  The
// array's name is a host-side stand-in: only this address may be passed to
// kernels and copies.
void* scratch_on(int dev) {
    if (dev < 0 || dev >= kMaxDevices) {
        mfx::set_error("device index out of range");
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_addr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            mfx::set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            (void)hipGetLastError();
            return nullptr;
        }
        g_addr[dev] = p;
    }
    return g_addr[dev];
}

}  // namespace

extern "C" {

int minfft_enqueue_noalloc(int kind, const void* in, void* out, const void* aux, void* stream) {
    mfx::clear_error();
    mfx::Plan* p = reinterpret_cast<mfx::Plan*>(const_cast<void*>(aux));
    if (!p || p->magic != 0x6d667870u) {
        mfx::set_error("invalid aux data");
        return -1;
    }
    if (kind < 0 || kind >= mfx::K_COUNT) {
        mfx::set_error("invalid kind");
        return -1;
    }
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != p->dev) {
        (void)hipGetLastError();
        mfx::set_error("minfft_enqueue_noalloc: the aux's device must be the current device");
        return -1;
    }
    if (!mfx::is_device_pointer(in) || !mfx::is_device_pointer(out)) {
        mfx::set_error("minfft_enqueue_noalloc needs device (or managed) pointers for in and out");
        return -1;
    }
    const size_t need = mfx::scratch_need(*p, kind, in, out);
    if (need > kScratchBytes) {
        mfx::set_error("minfft_enqueue_noalloc: this transform needs " + std::to_string(need) +
                       " bytes of scratch, more than the " + std::to_string(kScratchBytes) +
                       " bytes of the library's fixed scratch");
        return -1;
    }
    void* base = nullptr;
    if (need) {
        base = scratch_on(cur);
        if (!base) return -1;
    }
    hipStream_t st = stream ? reinterpret_cast<hipStream_t>(stream) : p->stream;
    return mfx::enqueue_scratch(*p, kind, in, out, st, base, kScratchBytes);
}

}  // extern "C"
