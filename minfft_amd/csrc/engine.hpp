// engine.hpp -- host side of the MI355X transform engine: plans ("aux"),
// pass scheduling and kernel launches.  Precision-generic; the two public
// C-ABI layers (FP64 and FP32 minfft.h) and the batch extension sit on top.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>
#include <vector>

#include "kinds.hpp"

namespace mfx {

enum Prec : int { PREC_F64 = 0, PREC_F32 = 1 };
enum Family : int { FAM_DFT = 0, FAM_REALDFT = 1, FAM_T2T3 = 2, FAM_T4 = 3 };

int family_of(int kind);

// Batch levels of one axis pass, slowest first; level 0 may use per-entry
// offset tables (device arrays) instead of a stride.
struct Levels {
    int n = 0;
    unsigned B[4] = {1, 1, 1, 1};
    long long is[4] = {0, 0, 0, 0};
    long long os[4] = {0, 0, 0, 0};
    const long long* ioff = nullptr;
    const long long* ooff = nullptr;
    const long long* foff = nullptr;   // two-step scratch offsets for the table level
    // complex pass over a real DFT's half spectrum whose last level indexes
    // h = 0..H-1 (stride 1), where a tile pass may transform columns 0 and
    // H-1 as one complex column: the forward pass right after the real pass
    // (both columns real), or the inverse pass right before it (only their
    // real parts are read)
    bool pack = false;
    // the pass is a whole 1D transform: its input is read once and its output
    // written once (non-temporal I/O where the kernel has it)
    bool stream_io = false;
    void push(unsigned b, long long i, long long o) {
        B[n] = b;
        is[n] = i;
        os[n] = o;
        ++n;
    }
    unsigned long long count() const {
        unsigned long long c = 1;
        for (int k = 0; k < n; ++k) c *= B[k];
        return c;
    }
};

struct Device {
    int id = -1;
    int lds_max = 65536;
    int ncu = 0;
    int nxcd = 8;
    int maxlog[2] = {0, 0};  // effective single-tile limit per precision
};

// Device resources shared by a plan's passes.
struct Workspace {
    void* tw = nullptr;        // table in the plan's precision
    void* tw64 = nullptr;      // FP64 master table
    int tw_log = 0;            // covers M <= 2^tw_log
    void* fbuf = nullptr;      // two-step intermediate
    void* gbuf = nullptr;      // two-step complex staging for real kinds
    size_t fs_bytes = 0;
    unsigned* cbuf = nullptr;  // counters of the XCD-cooperative kernel (batch plans)
    size_t cbuf_bytes = 0;
    int split16 = 0;           // FP32 data, radix-16 stages as split-FP16 MFMA products
};

// Launch schedule of the passes.  The defaults are the measured best for
// large batches on MI355X; MINFFT_TUNE="key=value,..." overrides them for a
// process, and batch plans pick one per group with the autotuner
// (autotune.cpp), installed for the group's launches with TuneScope.
struct Tune {
    int tpw = 256;      // target threads per work-group of row tiles
    int cfb = 128;      // bytes per row segment of a column (c-fast) tile
    int wave = 1;       // wavefront-local exchanges for row tiles of <= 64 lanes per transform
    int xcd = 1;        // XCD-aware tile order for column tiles
    int cfcore = -1;    // column tiles: 1 = c-fast core, 0 = row core + LDS-staged
                        // column I/O, -1 = by kind (complex 1, real 0)
    int stage = 0;      // force LDS-staged complex I/O
    int chunk_mb = 64;  // two-step chunk size for entry batches
    int split = 0;      // two-step split of a long axis: log2 N1 = log2(N)/2 + split
    int e32 = 0;        // radix-32 tiles (32 points per lane) where instantiated
    int split2 = -1;    // radix-2-split single-pass kernel: 1 = for the first length past
                        // one tile, 2 = also at the tile limit, 3 = also one below, 0 = off,
                        // -1 = by precision (FP32 3, FP64 2: the measured best)
    int fused = 1;      // fused two-step kernel for batches of long 1D entries
    int fused_wpc = 2;  // work-groups per CU of the fused kernel
    int fused_alt = 0;  // use the alternative (smaller) fused work-groups
    int e4 = 1;         // complex passes of at most 2^16 points in all (lengths
                        // 2^5..2^10) on 4-point-per-lane tiles: 1 = on, 0 = off
    int rsplit = 0;     // residue-split single-pass kernel for contiguous axes:
                        // 10*log2 P + residues per work-group (e.g. 31 = 8 blocks,
                        // one residue), where instantiated; 0 = off
    int coop = -1;      // XCD-cooperative form of the fused two-step kernel (coop.hpp):
                        // 1 on, 0 off, -1 by precision (FP64 on, FP32 off: the measured best)
    int coop_la = 8;    // its lookahead: rounds between a transform's two steps
    int mall_mb = 0;    // multi-dimensional passes on chunks of entries / slabs of
                        // about this many MiB, so the later passes read the MALL (0 = off:
                        // measured slower at every size, see run_entries)
    int mstages = 1;    // split-FP16 512-point tiles: radix-16 stages on the matrix cores
                        // (1: the first; 2: both -- measured slower, mfma.hpp)
    int rlong = 1;      // real/DCT/DST axes longer than a tile: 1 = two passes with the
                        // pre- and post-processing fused into the two-step FFT (rlong.hpp),
                        // 0 = pre-kernel, two-step FFT, post-kernel
    int csplit = -1;    // complex column passes (both sides columns of adjacent transforms)
                        // on the radix-2-split column tile (csplit.hpp) where instantiated:
                        // 1 = on, 0 = off (whole-transform tile), -1 = FP64 passes and FP32
                        // 512-point ones with strides of 1 MiB and more (the measured best)
};
struct Plan {
    unsigned magic = 0x6d667870u;
    Family fam = FAM_DFT;
    Prec prec = PREC_F64;
    int d = 0;
    std::vector<int> Ns;
    long long P = 0;          // product of dims
    int dev = -1;
    Workspace ws;
    void* tbuf = nullptr;     // full-array scratch ("t" of the reference aux)
    size_t tbuf_bytes = 0;
    void* stage = nullptr;    // staging buffer for host pointers (lazy)
    size_t stage_bytes = 0;
    void* hstage = nullptr;   // pinned host bounce buffer for host pointers (lazy,
    size_t hstage_bytes = 0;  // up to kPinnedStageMax bytes)
    unsigned* hflag = nullptr;  // completion word in coherent pinned memory (lazy)
    unsigned hseq = 0;          // last value queued for it
    std::vector<hipEvent_t> pev;  // completion of the pieces of a pipelined copy out (lazy)
    hipStream_t stream = nullptr;  // stream of synchronous calls and NULL-stream enqueues
    hipStream_t home = nullptr;    // the plan's own stream (made with the plan)
    Tune sched;               // launch schedule (autotuned for large plans)
    bool tuned = false;
};

// the schedule in force on this thread
const Tune& tune();
// applies "key=value,..." (the Tune fields: tpw, cfb, wave, xcd, cfcore,
// e32, e4, stage, chunk, split, split2, fused, wpc, falt, rsplit, coop,
// coopla, mall, rlong, mstages) to t; false on an unknown key
bool parse_tune(const char* spec, Tune& t);
// the process-wide schedule (defaults + MINFFT_TUNE)
const Tune& base_tune();
// installs `t` for the launches issued by this thread while alive
struct TuneScope {
    explicit TuneScope(const Tune* t);
    ~TuneScope();
    TuneScope(const TuneScope&) = delete;
    TuneScope& operator=(const TuneScope&) = delete;
    const Tune* prev;
};

// error channel (thread-local)
void set_error(const std::string& msg);
const char* last_error();
void clear_error();

Device& device();
bool ensure_device();

// Device allocation of the library (plans, scratch, staging, tuning
// buffers).  Test hook: with MINFFT_FAIL_ALLOC=k set, the k-th allocation
// after the variable took its current value, and every later one, fail as
// out of memory (the error paths of the makers and transforms).
hipError_t dev_alloc(void** p, size_t bytes);

size_t real_size(Prec p);
Plan* make_plan(Family fam, Prec prec, int d, const int* Ns);
void free_plan(Plan* p);

// sizes in bytes of the input / output of `kind` for this plan
size_t in_bytes(const Plan& p, int kind);
size_t out_bytes(const Plan& p, int kind);

// run `kind` on device pointers, asynchronously on `st`
int enqueue(Plan& p, int kind, const void* in, void* out, hipStream_t st);
// the same with the kernels' scratch (two-step intermediate, long real
// kinds' complex staging, the multi-dimensional real DFT's temp) carved
// from `scratch` (device memory of scratch_bytes, 256-byte aligned) instead
// of the plan's own buffers; -1 if it is too small
int enqueue_scratch(Plan& p, int kind, const void* in, void* out, hipStream_t st, void* scratch,
                    size_t scratch_bytes);
// bytes of scratch such a call needs (0: the transform uses none)
size_t scratch_need(const Plan& p, int kind, const void* in, const void* out);
// run `kind` with minfft semantics: any pointer kind, synchronous
int execute(Plan& p, int kind, const void* in, void* out);

// building blocks shared with the batch extension
int alloc_twiddles(Workspace& ws, Prec prec, int log_m, hipStream_t st);
void free_workspace(Workspace& ws);
int needed_tw_log(int kind, int logN);
// nullptr if a `kind` pass along an axis of 2^logN points fits the two-step
// decomposition at this precision, else the reason
const char* axis_length_error(Prec prec, int kind, int logN);
// one 1D pass of `kind` (length N = 2^logN along the axis) over a batch
int axis_pass(const Workspace& ws, Prec prec, int kind, int logN,
              const void* in, long long in_e, void* out, long long out_e,
              const Levels& lv, hipStream_t st);
// persistent launch for the short 1D complex groups of a batch (multi.hpp):
// resident work-groups on this device (0 = not available), and one launch
// (counter reset, then the kernel) on st
struct MultiArgs;
int multi_grid(Prec prec);
// queues (XCDs with an L2 of their own) of the XCD-cooperative two-step
// kernel on this device; 0 = the kernel is not used here
unsigned coop_queues();
// per_item: one work-group per item (grid = item count, no counter);
// otherwise `grid` persistent work-groups claim items from a.counter, which
// a one-lane kernel clears first
int launch_multi(Prec prec, bool inv, const MultiArgs& a, unsigned grid, bool per_item, hipStream_t st);
// schedule codes (Tune::rsplit) of the residue-split kernels instantiated
// for complex transforms of length 2^logN whose LDS fits this device
std::vector<int> rsplit_codes(Prec prec, int logN);
// true when a contiguous complex pass of length 2^logL runs as one
// single-pass launch under schedule t (tile, residue-split or radix-2-split
// kernel), i.e. touches no two-step scratch; the residue-split kernel runs
// only out of place
bool c2c_single_pass(Prec prec, int logL, const Tune& t, bool in_place = false);
// scratch (bytes) an axis pass of this shape needs in ws.fbuf / ws.gbuf;
// in_extent = number of input items addressed (for the mirrored layout)
size_t axis_scratch_bytes(Prec prec, int kind, int logN, unsigned long long ntrans,
                          unsigned long long in_extent);

bool is_device_pointer(const void* p);

// all passes of `kind` on `count` entries of dims Ns[0..d-1]; with tables
// (ioff/ooff: per-entry item offsets into in/out, foff: into the two-step
// scratch, toff: into tbuf) or, when they are null, a single transform at
// in/out.  tbuf: complex scratch of the leading-axes passes of the inverse
// real DFT (and of the aliased real DFT).
int run_entries(const Workspace& ws, Prec pr, int kind, int d, const int* Ns, const void* in, void* out,
                unsigned count, const long long* ioff, const long long* ooff, const long long* foff,
                void* tbuf, const long long* toff, hipStream_t st);

}  // namespace mfx
