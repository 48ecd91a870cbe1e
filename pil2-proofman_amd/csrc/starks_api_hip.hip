// Host side of libpil2_commit_hip.so: the C-ABI declared in include/pil2_commit_hip.h.
//
// Drop-in entry points keep the reference's names and argument meaning
// (pil2-stark/src/api/starks_api.hpp); the device-resident pil2hip_* entry points
// run the same operations on HBM buffers on a caller stream.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <map>
#include <mutex>
#include <set>
#include <string>
#include <vector>

#include "../../include/pil2_commit_hip.h"
#include "fri_kernels.h"
#include "gl_field.h"
#include "ntt_kernels.h"
#include "poseidon2_kernels.h"

// ------------------------------------------------------------------ state --
static std::mutex g_mu;
static uint8_t g_hash_family = 0;  // unset, as in the reference (hash_family.hpp:45-66)
static bool g_abort_on_error = true;
static thread_local std::string t_err;

static int64_t fail(int64_t code, const std::string &msg)
{
    t_err = msg;
    return code;
}

#define HIP_TRY(expr, what)                                                                               \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess) return fail(-100 - (int64_t)e_, std::string(what) + ": " + hipGetErrorString(e_)); \
    } while (0)

static const uint64_t kRoots[33] = {
    0x1ULL, 18446744069414584320ULL, 281474976710656ULL, 16777216ULL, 4096ULL, 64ULL, 8ULL,
    2198989700608ULL, 4404853092538523347ULL, 6434636298004421797ULL, 4255134452441852017ULL,
    9113133275150391358ULL, 4355325209153869931ULL, 4308460244895131701ULL,
    7126024226993609386ULL, 1873558160482552414ULL, 8167150655112846419ULL,
    5718075921287398682ULL, 3411401055030829696ULL, 8982441859486529725ULL,
    1971462654193939361ULL, 6553637399136210105ULL, 8124823329697072476ULL,
    5936499541590631774ULL, 2709866199236980323ULL, 8877499657461974390ULL,
    3757607247483852735ULL, 4969973714567017225ULL, 2147253751702802259ULL,
    2530564950562219707ULL, 1905180297017055339ULL, 3524815499551269279ULL,
    7277203076849721926ULL};  // goldilocks_base_field.cpp:11-44

// Per-device twiddle tables (built once per device, kept until process exit). They
// live in a device array of the code object rather than in hipMalloc'd memory, so
// that no entry point allocates for them: each device holds its own copy, found by
// hipGetSymbolAddress with that device current.
static const int kMaxDevices = 16;
// compact in-tile tables (2 x 2^15 words), then the split tables of the 2^24 domain
// (6 x 2^12 words) and of the 2^28 domain (6 x 2^14 words)
static const size_t kTabRoots = (size_t)2 << NTT_LOG_TILE, kTab24 = (size_t)1 << 12, kTab28 = (size_t)1 << 14;
static const size_t kTabWords = 2 * kTabRoots + 6 * kTab24 + 6 * kTab28;
static __device__ __align__(256) uint64_t g_tab_dev[kTabWords];
static NttTablesHost g_tables[kMaxDevices];
static bool g_tables_ready[kMaxDevices];

// Split tables of the 2^(2*bits) domain at t: w^x = lo[x & (2^bits - 1)] * hi[x >> bits]
// for w = W[2*bits] (forward and inverse), and 7^e the same way.
static void split_tables(uint64_t *t, uint32_t bits)
{
    const size_t nt = (size_t)1 << bits;
    uint64_t w = kRoots[2 * bits], wi = gl_inv(w);
    uint64_t *twl_f = t, *twh_f = t + nt, *twl_i = t + 2 * nt, *twh_i = t + 3 * nt, *shl = t + 4 * nt,
             *shh = t + 5 * nt;
    uint64_t wn = gl_pow(w, nt), win = gl_pow(wi, nt), sn = gl_pow(7, nt);
    uint64_t a = 1, b = 1, c = 1, d = 1, e = 1, f = 1;
    for (size_t j = 0; j < nt; j++) {
        twl_f[j] = a; twh_f[j] = b; twl_i[j] = c; twh_i[j] = d; shl[j] = e; shh[j] = f;
        a = gl_mul(a, w); b = gl_mul(b, wn);
        c = gl_mul(c, wi); d = gl_mul(d, win);
        e = gl_mul(e, 7); f = gl_mul(f, sn);
    }
}

static int64_t ensure_tables(NttTablesHost **out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(-4, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_tables_ready[dev]) {
        const size_t nr = kTabRoots;
        std::vector<uint64_t> h(kTabWords);
        uint64_t *rf = h.data(), *ri = rf + nr, *t24 = ri + nr, *t28 = t24 + 6 * kTab24;
        for (uint32_t k = 0; k <= NTT_LOG_TILE; k++) {  // rf[2^k + y] = w_{2^k}^y, ri[2^k + y] = w_{2^k}^-y
            uint64_t w = kRoots[k], wi = gl_inv(w), a = 1, b = 1;
            for (size_t y = 0; y < ((size_t)1 << k); y++) {
                rf[((size_t)1 << k) + y] = a;
                ri[((size_t)1 << k) + y] = b;
                a = gl_mul(a, w);
                b = gl_mul(b, wi);
            }
        }
        rf[0] = ri[0] = 1;  // unused slot
        int ie = ntt_init_device();
        if (ie) return fail(-100 - ie, "NTT device init failed");
        split_tables(t24, 12);
        split_tables(t28, 14);
        uint64_t *dptr = nullptr;
        HIP_TRY(hipGetSymbolAddress((void **)&dptr, HIP_SYMBOL(g_tab_dev)), "hipGetSymbolAddress(tables)");
        HIP_TRY(hipMemcpy(dptr, h.data(), h.size() * 8, hipMemcpyHostToDevice), "hipMemcpy(tables)");
        NttTablesHost &T = g_tables[dev];
        T.tkf = dptr;
        T.tki = dptr + nr;
        for (int d = 0; d < 2; d++) {
            size_t nt = d ? kTab28 : kTab24;
            uint64_t *dt = dptr + 2 * nr + (d ? 6 * kTab24 : 0);
            NttSplitTables &S = T.split[d];
            S.twl_f = dt; S.twh_f = dt + nt; S.twl_i = dt + 2 * nt; S.twh_i = dt + 3 * nt; S.shl = dt + 4 * nt;
            S.shh = dt + 5 * nt;
            S.bits = d ? 14 : 12;
        }
        g_tables_ready[dev] = true;
    }
    *out = &g_tables[dev];
    return 0;
}

// The root tables for the FRI folds. Setting them up copies them to the device
// synchronously, which a stream capture does not allow: on a stream that is not
// being captured they are set up here if needed; under capture only tables set
// up earlier are used, and without them the folds raise the root to a power.
static const NttTablesHost *fri_tables(hipStream_t st)
{
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess) {
        (void)hipGetLastError();
        cs = hipStreamCaptureStatusActive;
    }
    NttTablesHost *tb = nullptr;
    if (cs == hipStreamCaptureStatusNone) return ensure_tables(&tb) ? nullptr : tb;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_mu);
    return g_tables_ready[dev] ? &g_tables[dev] : nullptr;
}

// ----------------------------------------------------------- phase timing --
// When enabled, pil2hip_commit records HIP events on the caller's stream
// between its kernels and accumulates the elapsed time per phase:
// 0 transpose, 1 iNTT strided pass, 2 fused LDE block, 3 NTT strided pass,
// 4 leaf hashing, 5 tree levels. Timed commits synchronise their stream.
static const int kPhases = 6;
static bool g_timing = false;
static bool g_timing_ev_ready = false;
static hipEvent_t g_ev[kPhases + 1];
static double g_phase_ms[kPhases];
static uint64_t g_timed_calls = 0;

static thread_local bool t_capturing = false;  // no timing events inside graph capture

static hipEvent_t *timing_events()
{
    if (!g_timing || t_capturing) return nullptr;
    if (!g_timing_ev_ready) {
        for (int i = 0; i <= kPhases; i++)
            if (hipEventCreate(&g_ev[i]) != hipSuccess) return nullptr;
        g_timing_ev_ready = true;
    }
    return g_ev + 1;  // ntt_launch_lde records ev[1..4]; ev[0] is the start
}

static int64_t check_family()
{
    if (g_hash_family == 2) return 0;
    if (g_hash_family == 1) return fail(-5, "hash family Poseidon1 is not implemented by this library (Poseidon2 only)");
    return fail(-6, "hash family not set: call set_hash_family(2) first");
}

static int64_t check_arity(uint64_t arity)
{
    if (arity < 2 || arity > 4) return fail(-7, "unsupported arity (Poseidon2 supports 2, 3, 4)");
    return 0;
}

// measured on the plain workloads: 2^13 strided tiles (two blocks per CU) and radix-16
// block rounds (plain-large 118.9 ms against 119.4 ms with radix-8 rounds, now that blowup
// 4 runs two columns per block; plain-small the same either way)
static const Pil2CommitPlan kDefaultPlan = {1, 13, 4, P2_COOP_MAX_PARENTS};

static int64_t lde_impl(uint64_t *d_lde, const uint64_t *d_trace, uint64_t *d_scratch, uint64_t nBits,
                        uint64_t nBitsExt, uint64_t nCols, hipStream_t st, hipEvent_t *ev,
                        const Pil2CommitPlan &plan = kDefaultPlan)
{
    if (nBitsExt < nBits) return fail(-10, "nBitsExt < nBits");
    if (nBitsExt > NTT_MAX_LOG) return fail(-11, "nBitsExt > 28 is not supported");
    if (nCols == 0) return 0;
    if (nCols > 65535) return fail(-12, "nCols > 65535 is not supported");
    if (!d_lde || !d_trace || !d_scratch) return fail(-13, "null buffer");
    NttTablesHost *tb;
    int64_t rc = ensure_tables(&tb);
    if (rc) return rc;
    NttPlan np = {plan.ntt_tile_log, plan.ntt_radix_log};
    int e = ntt_launch_lde(d_lde, d_trace, d_scratch, (uint32_t)nBits, (uint32_t)nBitsExt, nCols, tb, &np, st, ev);
    if (e < 0) return fail(-14, "LDE plan not supported for this shape");
    if (e) return fail(-100 - e, std::string("LDE launch: ") + hipGetErrorString((hipError_t)e));
    return 0;
}

// LDE + leaves + tree levels on stream st. With ev (the phase events of
// g_ev[1..]), the phases are timed and the stream is synchronised.
static int64_t commit_impl(uint64_t *d_lde, uint64_t *d_nodes, const uint64_t *d_trace, uint64_t *d_scratch,
                           uint64_t arity, uint64_t nBits, uint64_t nBitsExt, uint64_t nCols, hipStream_t st,
                           hipEvent_t *ev, const Pil2CommitPlan &plan = kDefaultPlan)
{
    int64_t rc = check_family();
    if (rc) return rc;
    if ((rc = check_arity(arity))) return rc;
    if (nBitsExt > NTT_MAX_LOG) return fail(-11, "nBitsExt > 28 is not supported");
    if (!d_nodes) return fail(-13, "null buffer");
    if (ev) (void)hipEventRecord(g_ev[0], st);
    if ((rc = lde_impl(d_lde, d_trace, d_scratch, nBits, nBitsExt, nCols, st, ev, plan))) return rc;
    uint64_t NExt = 1ull << nBitsExt;
    uint32_t width = (uint32_t)(4 * arity);
    int e = p2_launch_leaves(width, d_lde, 1, NExt, nCols, NExt, d_nodes, st);
    if (e) return fail(-100 - e, "leaf hashing launch failed");
    if (ev) (void)hipEventRecord(g_ev[5], st);
    e = p2_launch_tree_levels(width, d_nodes, NExt, plan.coop_max_parents, st);
    if (e) return fail(-100 - e, "tree level launch failed");
    if (ev) {
        HIP_TRY(hipEventRecord(g_ev[6], st), "timing event");
        HIP_TRY(hipEventSynchronize(g_ev[6]), "timing sync");
        for (int i = 0; i < kPhases; i++) {
            float ms = 0;
            if (hipEventElapsedTime(&ms, g_ev[i], g_ev[i + 1]) == hipSuccess) g_phase_ms[i] += ms;
        }
        g_timed_calls++;
    }
    return 0;
}

// Internal interface for the other host sources of the library (no timing).
int64_t commit_enqueue(uint64_t *d_lde, uint64_t *d_nodes, const uint64_t *d_trace, uint64_t *d_scratch,
                       uint64_t arity, uint64_t nBits, uint64_t nBitsExt, uint64_t nCols, hipStream_t st)
{
    return commit_impl(d_lde, d_nodes, d_trace, d_scratch, arity, nBits, nBitsExt, nCols, st, nullptr);
}

int64_t api_fail(int64_t code, const std::string &msg) { return fail(code, msg); }

// ---------------------------------------------------------- entry points --
extern "C" {

void set_hash_family(uint8_t fam)
{
    std::lock_guard<std::mutex> lk(g_mu);
    g_hash_family = fam;
}

void pil2hip_set_abort_on_error(bool v) { g_abort_on_error = v; }

const char *pil2hip_last_error(void) { return t_err.c_str(); }

uint64_t get_num_gpus(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return (uint64_t)n;
}

bool set_gpu_mode(bool use_gpu) { return use_gpu; }

struct DeviceBuffersHip {
    int device;
    hipStream_t stream;
    uint32_t arity;
    uint32_t max_n_bits_ext;
};

void *gen_device_buffers(uint32_t node_rank, uint32_t node_size, const int32_t *numa_nodes, uint32_t arity,
                         uint32_t max_n_bits_ext)
{
    (void)node_rank; (void)node_size; (void)numa_nodes;
    DeviceBuffersHip *b = new DeviceBuffersHip();
    if (hipGetDevice(&b->device) != hipSuccess || hipStreamCreate(&b->stream) != hipSuccess) {
        delete b;
        fail(-8, "gen_device_buffers: no HIP device");
        return nullptr;
    }
    b->arity = arity;
    b->max_n_bits_ext = max_n_bits_ext;
    return b;
}

void free_device_buffers(void *d_buffers)
{
    DeviceBuffersHip *b = (DeviceBuffersHip *)d_buffers;
    if (!b) return;
    (void)hipStreamSynchronize(b->stream);
    (void)hipStreamDestroy(b->stream);
    delete b;
}

uint64_t pil2hip_tree_num_elements(uint64_t rows, uint64_t arity)
{
    if (arity < 2) return 0;
    uint64_t total = rows, level = rows;
    while (level > 1) {
        total += (arity - level % arity) % arity;
        level = (level + arity - 1) / arity;
        total += level;
    }
    return total * 4;
}

int64_t pil2hip_lde(uint64_t *d_lde, const uint64_t *d_trace, uint64_t *d_scratch, uint64_t nBits, uint64_t nBitsExt,
                    uint64_t nCols, void *stream)
{
    return lde_impl(d_lde, d_trace, d_scratch, nBits, nBitsExt, nCols, (hipStream_t)stream, timing_events());
}

int64_t pil2hip_merkelize(uint64_t *d_nodes, const uint64_t *d_src, uint64_t nRows, uint64_t nCols, uint64_t rowStride,
                          uint64_t colStride, uint64_t arity, void *stream)
{
    int64_t rc = check_family();
    if (rc) return rc;
    if ((rc = check_arity(arity))) return rc;
    if (nRows == 0) return 0;
    if (!d_nodes || (!d_src && nCols)) return fail(-13, "null buffer");
    uint32_t width = (uint32_t)(4 * arity);
    int e = p2_launch_leaves(width, d_src, rowStride, colStride, nCols, nRows, d_nodes, (hipStream_t)stream);
    if (e) return fail(-100 - e, "leaf hashing launch failed");
    e = p2_launch_tree_levels(width, d_nodes, nRows, P2_COOP_MAX_PARENTS, (hipStream_t)stream);
    if (e) return fail(-100 - e, "tree level launch failed");
    return 0;
}

int64_t pil2hip_commit(uint64_t *d_lde, uint64_t *d_nodes, const uint64_t *d_trace, uint64_t *d_scratch, uint64_t arity,
                       uint64_t nBits, uint64_t nBitsExt, uint64_t nCols, void *stream)
{
    return commit_impl(d_lde, d_nodes, d_trace, d_scratch, arity, nBits, nBitsExt, nCols, (hipStream_t)stream,
                       timing_events());
}

// ------------------------------------------------- no-allocation commit --
// pil2hip_commit for callers that cannot allocate (graph capture, tight serving
// loops): the column-major trace scratch is this file-scope device array, of
// PIL2HIP_NOALLOC_SCRATCH_WORDS words per device, held until the process ends.
// In host code `scratch` names a host-side stand-in, never a device address: the
// address of each device's copy comes from hipGetSymbolAddress with that device
// current, once per device, and only that pointer reaches kernels.
static __device__ __align__(256) uint64_t scratch[PIL2HIP_NOALLOC_SCRATCH_WORDS];
static uint64_t *g_scratch_ptr[kMaxDevices];

static int64_t noalloc_scratch(uint64_t **out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(-4, "device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch_ptr[dev]) {
        uint64_t *p = nullptr;
        HIP_TRY(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

int64_t pil2hip_commit_noalloc(uint64_t *d_lde, uint64_t *d_nodes, const uint64_t *d_trace, uint64_t arity,
                               uint64_t nBits, uint64_t nBitsExt, uint64_t nCols, void *stream)
{
    if (nBits > 63 || (nCols && ((1ull << nBits) > PIL2HIP_NOALLOC_SCRATCH_WORDS / nCols)))
        return fail(-36, "pil2hip_commit_noalloc: 2^nBits * nCols exceeds PIL2HIP_NOALLOC_SCRATCH_WORDS");
    // the twiddle tables (a device array too) and the scratch address, once per device
    NttTablesHost *tb;
    int64_t rc = ensure_tables(&tb);
    if (rc) return rc;
    uint64_t *s = nullptr;
    if ((rc = noalloc_scratch(&s))) return rc;
    return commit_impl(d_lde, d_nodes, d_trace, s, arity, nBits, nBitsExt, nCols, (hipStream_t)stream, nullptr);
}

// ------------------------------------------------------- streamed commit --
// A host-resident trace committed with its transfer overlapped: the column-chunk
// absorb of the reference's streaming commit (stream_commit.cuh:9-20), here for
// an unpacked trace and Poseidon2. Columns go in groups of w (a whole number of
// rate blocks, except the last group). A copy stream moves group g out of the
// host rows (hipMemcpy2DAsync) into one of two row-major staging buffers while
// the compute stream extends group g-1 into its columns of the LDE and absorbs
// them into the row sponges, whose capacity carries over between groups in the
// leaf-digest words of d_nodes. The tree levels follow the last group.
// Automatic width: one group for traces under 256 MiB (the copy is short and
// per-group launches would cost more than it hides); otherwise about eight
// groups of whole rate blocks, at least 16 columns wide: host rows narrower than
// 128 B made the 2D copy about half as fast (2^22 x 128 at blowup 4 from pinned
// memory: 8 columns 231.6 ms, 16 columns 150.5 ms, 32 columns 152.1 ms, 64
// columns 158.1 ms).
static uint64_t stream_group_cols(uint64_t nBits, uint64_t nCols, uint64_t arity, uint64_t groupCols)
{
    uint64_t all = std::max<uint64_t>(nCols, 1);
    if (groupCols) return std::min(groupCols, all);
    if (nBits >= 40 || (nCols << nBits) * 8 < (256ull << 20)) return all;
    uint64_t rate = 4 * arity - 4, blocks = (nCols + rate - 1) / rate;
    uint64_t w = std::max(rate * ((16 + rate - 1) / rate), rate * ((blocks + 7) / 8));
    return std::min(w, all);
}

// Width of the groups after the first. With the automatic width the first group
// stays narrow (it is copied before any compute can start) and the later ones
// are twice as wide: a 2D copy of 2^22 rows costs about the same per row piece
// up to 128 bytes, so 256-byte pieces move 1.4x the bytes per second (43.7 against
// 30.6 GB/s) and the copies then keep ahead of the compute.
static uint64_t stream_later_cols(uint64_t nBits, uint64_t nCols, uint64_t arity, uint64_t groupCols)
{
    uint64_t w = stream_group_cols(nBits, nCols, arity, groupCols);
    if (groupCols || w >= nCols) return w;
    return std::min(2 * w, nCols - w);
}

uint64_t pil2hip_commit_host_stage_words(uint64_t nBits, uint64_t nCols, uint64_t arity, uint64_t groupCols)
{
    if (arity < 2 || arity > 4 || nBits > 40) return 0;
    uint64_t w = std::max(stream_group_cols(nBits, nCols, arity, groupCols),
                          stream_later_cols(nBits, nCols, arity, groupCols));
    return 3 * (1ull << nBits) * w;
}

int64_t pil2hip_commit_host(uint64_t *d_lde, uint64_t *d_nodes, const uint64_t *h_trace, uint64_t *d_stage,
                            uint64_t arity, uint64_t nBits, uint64_t nBitsExt, uint64_t nCols, uint64_t groupCols,
                            void *stream)
{
    int64_t rc = check_family();
    if (rc) return rc;
    if ((rc = check_arity(arity))) return rc;
    if (nBitsExt < nBits) return fail(-10, "nBitsExt < nBits");
    if (nBitsExt > NTT_MAX_LOG) return fail(-11, "nBitsExt > 28 is not supported");
    if (nCols > 65535) return fail(-12, "nCols > 65535 is not supported");
    if (!d_nodes || (nCols && (!d_lde || !h_trace || !d_stage))) return fail(-13, "null buffer");
    uint64_t rate = 4 * arity - 4, w = stream_group_cols(nBits, nCols, arity, groupCols);
    if (w < nCols && w % rate) return fail(-33, "groupCols must be a multiple of the sponge rate 4*arity-4");
    if (nCols) {
        // From pageable memory the runtime stages a 2D copy in small pieces and a
        // 2^22-row group takes minutes, so only page-locked host memory is taken.
        hipPointerAttribute_t pa;
        hipError_t pe = hipPointerGetAttributes(&pa, h_trace);
        (void)hipGetLastError();  // an unregistered pointer leaves an error that the launches below would report
        if (pe != hipSuccess || pa.type != hipMemoryTypeHost)
            return fail(-34, "h_trace must be page-locked host memory (hipHostMalloc)");
    }
    hipStream_t st = (hipStream_t)stream;
    uint64_t N = 1ull << nBits, NExt = 1ull << nBitsExt;
    uint32_t width = (uint32_t)(4 * arity);
    int e = 0;
    if (nCols == 0) {  // empty rows: zero digests
        e = p2_launch_leaves(width, d_lde, 1, NExt, 0, NExt, d_nodes, st);
    } else if (w >= nCols) {  // one group: a plain copy on the caller's stream, nothing to overlap
        hipError_t he = hipMemcpyAsync(d_stage, h_trace, N * nCols * 8, hipMemcpyHostToDevice, st);
        if (he == hipSuccess) he = hipStreamSynchronize(st);  // h_trace is free on return
        if (he != hipSuccess) return fail(-100 - (int64_t)he, "streamed commit: copy failed");
        if ((rc = lde_impl(d_lde, d_stage, d_stage + 2 * N * w, nBits, nBitsExt, nCols, st, nullptr))) return rc;
        e = p2_launch_leaves(width, d_lde, 1, NExt, nCols, NExt, d_nodes, st);
    } else {
        hipStream_t cs = nullptr;
        hipEvent_t copied[2] = {nullptr, nullptr}, freed[2] = {nullptr, nullptr}, gathered = nullptr;
        auto done = [&](int64_t r) {
            // the host rows are free after this; d_stage stays in use until `st` has run the last group's LDE
            if (cs) (void)hipStreamSynchronize(cs);
            if (gathered) (void)hipEventSynchronize(gathered);  // the kernel that read host rows on `st`
            if (r) (void)hipStreamSynchronize(st);   // on an error nothing queued may still read d_stage
            for (int k = 0; k < 2; k++) {
                if (copied[k]) (void)hipEventDestroy(copied[k]);
                if (freed[k]) (void)hipEventDestroy(freed[k]);
            }
            if (gathered) (void)hipEventDestroy(gathered);
            if (cs) (void)hipStreamDestroy(cs);
            return r;
        };
        hipError_t he = hipStreamCreateWithFlags(&cs, hipStreamNonBlocking);
        for (int k = 0; k < 2 && he == hipSuccess; k++) {
            he = hipEventCreateWithFlags(&copied[k], hipEventDisableTiming);
            if (he == hipSuccess) he = hipEventCreateWithFlags(&freed[k], hipEventDisableTiming);
        }
        if (he == hipSuccess) he = hipEventCreateWithFlags(&gathered, hipEventDisableTiming);
        if (he != hipSuccess) return done(fail(-100 - (int64_t)he, "streamed commit: stream/event creation failed"));
        // The first group's copy cannot overlap any compute (its LDE is the first
        // work), so a kernel on `st`, with the device otherwise idle, reads those
        // page-locked rows itself: 42 GB/s for 128-byte pieces against about 30 for
        // the 2D DMA copy, while the DMA engine starts on the second group at once.
        // (Giving the DMA engine a share of the first group was slower: both read
        // over the same link.) Without a device pointer for h_trace the DMA copies it.
        void *hd = nullptr;
        bool gather = hipHostGetDevicePointer(&hd, (void *)h_trace, 0) == hipSuccess && hd;
        (void)hipGetLastError();
        uint64_t r_dma = gather ? 0 : N;
        uint64_t wl = stream_later_cols(nBits, nCols, arity, groupCols), wmax = std::max(w, wl);
        uint64_t *scratch = d_stage + 2 * N * wmax;
        for (uint64_t g = 0, c0 = 0; c0 < nCols; c0 += (g == 0 ? w : wl), g++) {
            uint64_t wg = std::min(g == 0 ? w : wl, nCols - c0), *buf = d_stage + (g & 1) * N * wmax;
            uint64_t rows = g == 0 ? r_dma : N;
            if (g >= 2) he = hipStreamWaitEvent(cs, freed[g & 1], 0);  // group g-2's LDE has read buf
            if (he == hipSuccess && rows)
                he = hipMemcpy2DAsync(buf, wg * 8, h_trace + c0, nCols * 8, wg * 8, rows, hipMemcpyHostToDevice, cs);
            if (he == hipSuccess && g == 0 && rows < N) {
                e = ntt_launch_gather_rows(buf + rows * wg, (const uint64_t *)hd + rows * nCols + c0, N - rows,
                                           (uint32_t)wg, nCols, st);
                if (e) return done(fail(-100 - e, "streamed commit: host-row gather launch failed"));
                he = hipEventRecord(gathered, st);
            }
            if (he == hipSuccess) he = hipEventRecord(copied[g & 1], cs);
            if (he == hipSuccess) he = hipStreamWaitEvent(st, copied[g & 1], 0);
            if (he != hipSuccess) return done(fail(-100 - (int64_t)he, "streamed commit: copy failed"));
            if ((rc = lde_impl(d_lde + c0 * NExt, buf, scratch, nBits, nBitsExt, wg, st, nullptr))) return done(rc);
            if ((he = hipEventRecord(freed[g & 1], st)) != hipSuccess)
                return done(fail(-100 - (int64_t)he, "streamed commit: event record failed"));
            e = p2_launch_leaves_carry(width, d_lde + c0 * NExt, 1, NExt, wg, NExt, d_nodes, g > 0, st);
            if (e) return done(fail(-100 - e, "leaf hashing launch failed"));
        }
        done(0);
    }
    if (e) return fail(-100 - e, "leaf hashing launch failed");
    e = p2_launch_tree_levels(width, d_nodes, NExt, P2_COOP_MAX_PARENTS, st);
    if (e) return fail(-100 - e, "tree level launch failed");
    return 0;
}

int64_t pil2hip_ntt(uint64_t *d_dst, const uint64_t *d_src, uint64_t *d_scratch, uint64_t nBits, uint64_t nCols,
                    bool inverse, bool rowMajor, void *stream)
{
    if (nBits > NTT_MAX_LOG) return fail(-11, "size > 2^28 is not supported");
    if (nCols == 0) return 0;
    if (nCols > 65535) return fail(-12, "nCols > 65535 is not supported");
    if (!d_src) return fail(-13, "null buffer");
    if (!d_dst) d_dst = (uint64_t *)d_src;  // NTT_Goldilocks: dst == NULL transforms in place
    uint64_t M = 1ull << nBits;
    hipStream_t st = (hipStream_t)stream;
    int e;
    if (nBits < 4) {
        e = ntt_launch_ntt_tiny(d_dst, d_src, rowMajor ? nCols : 1, rowMajor ? 1 : M, (uint32_t)nBits, nCols, inverse,
                                st);
    } else {
        NttTablesHost *tb;
        int64_t rc = ensure_tables(&tb);
        if (rc) return rc;
        if (rowMajor) {  // rows -> columns, transform out of place, columns -> rows
            if (!d_scratch) return fail(-13, "row-major transforms need a scratch of 2 * 2^nBits * nCols words");
            uint64_t *a = d_scratch, *b = d_scratch + M * nCols;
            e = ntt_launch_rm2cm(a, d_src, M, nCols, nCols, M, st);
            if (!e) e = ntt_launch_ntt(b, a, (uint32_t)nBits, nCols, inverse, tb, st);
            if (!e) e = ntt_launch_cm2rm(d_dst, b, M, nCols, M, nCols, st);
        } else if (d_dst == d_src) {
            if (!d_scratch) return fail(-13, "in-place column-major transforms need a scratch of 2^nBits * nCols words");
            e = ntt_launch_ntt(d_scratch, d_src, (uint32_t)nBits, nCols, inverse, tb, st);
            if (!e) e = (int)hipMemcpyAsync(d_dst, d_scratch, M * nCols * 8, hipMemcpyDeviceToDevice, st);
        } else {
            e = ntt_launch_ntt(d_dst, d_src, (uint32_t)nBits, nCols, inverse, tb, st);
        }
    }
    if (e < 0) return fail(-14, "NTT plan not supported for this shape");
    if (e) return fail(-100 - e, std::string("NTT launch: ") + hipGetErrorString((hipError_t)e));
    return 0;
}

int64_t pil2hip_compute_q(uint64_t *d_cmQ, uint64_t *d_nodes, const uint64_t *d_q, uint64_t qRowStride,
                          uint64_t qColStride, uint64_t *d_scratch, uint64_t nBits, uint64_t nBitsExt, uint64_t qDeg,
                          uint64_t qDim, uint64_t arity, void *stream)
{
    int64_t rc;
    if (d_nodes) {
        if ((rc = check_family())) return rc;
        if ((rc = check_arity(arity))) return rc;
    }
    if (nBitsExt < nBits) return fail(-10, "nBitsExt < nBits");
    if (nBitsExt > NTT_MAX_LOG) return fail(-11, "nBitsExt > 28 is not supported");
    if (qDeg == 0 || qDim == 0) return fail(-12, "qDeg and qDim must be >= 1");
    if (qDeg > (1ull << (nBitsExt - nBits))) return fail(-12, "qDeg exceeds the blowup factor NExt / N");
    if (qDeg * qDim > 65535) return fail(-12, "qDeg * qDim > 65535 is not supported");
    uint64_t NExt = 1ull << nBitsExt;
    if (qRowStride == 1 ? (qDim > 1 && qColStride < NExt) : (qColStride != 1 || qRowStride < qDim))
        return fail(-16, "q must be column-major (rowStride 1, colStride >= NExt) or row-major (colStride 1)");
    if (!d_cmQ || !d_q || !d_scratch) return fail(-13, "null buffer");
    NttTablesHost *tb;
    if ((rc = ensure_tables(&tb))) return rc;
    hipStream_t st = (hipStream_t)stream;
    int e = ntt_launch_compute_q(d_cmQ, d_q, qRowStride, qRowStride == 1 && qDim == 1 ? NExt : qColStride, d_scratch,
                                 (uint32_t)nBits, (uint32_t)nBitsExt, (uint32_t)qDeg, (uint32_t)qDim, tb, st);
    if (e < 0) return fail(-14, "computeQ plan not supported for this shape");
    if (e) return fail(-100 - e, std::string("computeQ launch: ") + hipGetErrorString((hipError_t)e));
    if (!d_nodes) return 0;
    return pil2hip_merkelize(d_nodes, d_cmQ, NExt, qDeg * qDim, 1, NExt, arity, stream);
}

// ------------------------------------------------------------ batched commit --
// Many independent commits (uneven shapes) in one hipGraph: every job's
// LDE + Merkle chain is enqueued on one of S internal streams (longest jobs
// first, each to the least-loaded stream), a one-block kernel copies each root
// into d_roots, and the whole fork/join capture is instantiated once per job
// list and plan and replayed. A branch count of 0 is tuned on the first call
// (S in {1, 2, 4, 8, 16}, each candidate timed over one replay of the whole
// batch after one warm replay). A plan with every field 0 also takes each job's
// NTT tile, radix and level threshold from the shape wisdom below (measured on
// first sight of a shape); otherwise plan fields left 0 take their defaults
// (tile 2^13, radix 16, cooperative levels up to 16,384 parents).
struct BatchGraph {
    std::vector<uint64_t> key;
    Pil2CommitPlan plan;
    hipGraphExec_t exec;
    hipEvent_t done;  // recorded after every launch of exec; waited on before exec is destroyed
};

// A cached executable graph may still run on the caller's stream when it is
// evicted or released: record an event after every launch and wait for it
// before the exec is destroyed. A launch into a stream that is being captured
// records no event (the caller's graph then holds its own copy of the work).
static hipError_t launch_cached(hipGraphExec_t ex, hipEvent_t done, hipStream_t st)
{
    hipError_t e = hipGraphLaunch(ex, st);
    if (e != hipSuccess) return e;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs == hipStreamCaptureStatusNone)
        e = hipEventRecord(done, st);
    return e;
}

static void retire_exec(hipGraphExec_t ex, hipEvent_t done)
{
    if (done) {
        (void)hipEventSynchronize(done);
        (void)hipEventDestroy(done);
    }
    if (ex) (void)hipGraphExecDestroy(ex);
}
static std::vector<BatchGraph> g_batch_graphs;
static std::vector<hipStream_t> g_batch_streams;
static std::vector<hipEvent_t> g_batch_events;
static hipStream_t g_capture_stream = nullptr;
static const uint32_t kMaxBatchStreams = 16;
static std::mutex g_batch_mu;  // batch graphs, streams and events are shared state

static double job_work(const Pil2CommitJob &j)
{
    // LDE element count weighted by the sponge permutations per row
    double rate = (double)(4 * j.arity - 4);
    return (double)(1ull << j.n_bits_ext) * (double)j.n_cols * (1.0 + 8.0 / rate);
}

static int64_t batch_resources()
{
    if (g_capture_stream) return 0;
    HIP_TRY(hipStreamCreateWithFlags(&g_capture_stream, hipStreamNonBlocking), "capture stream");
    g_batch_streams.resize(kMaxBatchStreams);
    g_batch_events.resize(kMaxBatchStreams + 1);
    for (uint32_t i = 0; i < kMaxBatchStreams; i++)
        HIP_TRY(hipStreamCreateWithFlags(&g_batch_streams[i], hipStreamNonBlocking), "batch stream");
    for (auto &e : g_batch_events) HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming), "batch event");
    return 0;
}

static Pil2CommitPlan job_plan(const Pil2CommitJob &j, const Pil2CommitPlan &plan);

static int64_t batch_capture(const Pil2CommitJob *jobs, uint64_t n, const Pil2CommitPlan &plan, uint64_t *d_roots,
                             hipGraphExec_t *out)
{
    // LPT assignment: jobs by decreasing work, each to the least-loaded stream
    uint32_t S = plan.n_streams;
    std::vector<uint64_t> order(n);
    for (uint64_t i = 0; i < n; i++) order[i] = i;
    std::sort(order.begin(), order.end(), [&](uint64_t a, uint64_t b) { return job_work(jobs[a]) > job_work(jobs[b]); });
    std::vector<double> load(S, 0.0);
    std::vector<uint32_t> assign(n);
    for (uint64_t k = 0; k < n; k++) {
        uint32_t best = 0;
        for (uint32_t s = 1; s < S; s++)
            if (load[s] < load[best]) best = s;
        assign[order[k]] = best;
        load[best] += job_work(jobs[order[k]]);
    }
    hipStream_t cap = g_capture_stream;
    HIP_TRY(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal), "begin capture");
    t_capturing = true;
    int64_t rc = 0;
    hipError_t fe = hipEventRecord(g_batch_events[kMaxBatchStreams], cap);
    for (uint32_t s = 0; s < S && fe == hipSuccess; s++)
        fe = hipStreamWaitEvent(g_batch_streams[s], g_batch_events[kMaxBatchStreams], 0);
    if (fe != hipSuccess) rc = fail(-100 - (int64_t)fe, std::string("graph fork: ") + hipGetErrorString(fe));
    for (uint64_t k = 0; k < n && !rc; k++) {
        const Pil2CommitJob &j = jobs[order[k]];
        hipStream_t st = g_batch_streams[assign[order[k]]];
        rc = commit_impl(j.d_lde, j.d_nodes, j.d_trace, j.d_scratch, j.arity, j.n_bits, j.n_bits_ext, j.n_cols, st,
                         nullptr, job_plan(j, plan));
        if (!rc && d_roots) {
            uint64_t nt = pil2hip_tree_num_elements(1ull << j.n_bits_ext, j.arity);
            int e = p2_launch_copy_root(j.d_nodes + nt - 4, d_roots + 4 * order[k], st);
            if (e) rc = fail(-100 - e, "root copy launch failed");
        }
    }
    for (uint32_t s = 0; s < S; s++) {  // join every branch (also after an error, to end the capture cleanly)
        hipError_t je = hipEventRecord(g_batch_events[s], g_batch_streams[s]);
        if (je == hipSuccess) je = hipStreamWaitEvent(cap, g_batch_events[s], 0);
        if (je != hipSuccess && !rc) rc = fail(-100 - (int64_t)je, std::string("graph join: ") + hipGetErrorString(je));
    }
    t_capturing = false;
    hipGraph_t graph = nullptr;
    hipError_t ee = hipStreamEndCapture(cap, &graph);
    if (rc) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (ee != hipSuccess) return fail(-100 - (int64_t)ee, std::string("end capture: ") + hipGetErrorString(ee));
    ee = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ee != hipSuccess) return fail(-100 - (int64_t)ee, std::string("graph instantiate: ") + hipGetErrorString(ee));
    return 0;
}

// ---- shape wisdom -------------------------------------------------------
// The batch autotuner's per-shape choices: the NTT block tile and in-tile radix
// per LDE shape (nBits, nBitsExt, ceil(log2 nCols)), and the cooperative-level
// threshold per tree shape (nBitsExt, arity). Each shape is measured once per
// process, eagerly (outside any graph) on the buffers of the first batch job that
// has it, and every later batch reuses the entry. Plans whose launches coincide
// for a shape (ntt_lde_plan_signature, level_schedule) are timed once.
static std::map<uint64_t, NttPlan> g_lde_wisdom;
static std::map<uint64_t, uint32_t> g_level_wisdom;
static const uint32_t kLevelThresholds[] = {4096, 16384, 65536};

static uint64_t lde_key(const Pil2CommitJob &j)
{
    uint64_t lc = 0;
    while ((1ull << lc) < j.n_cols) lc++;
    return j.n_bits | j.n_bits_ext << 8 | lc << 16;
}
static uint64_t level_key(const Pil2CommitJob &j) { return j.n_bits_ext | j.arity << 8; }

// Number of tree levels a threshold sends to the cooperative kernel (the level
// schedule p2_launch_tree_levels follows).
static uint32_t level_schedule(uint64_t nleaves, uint64_t arity, uint64_t thr)
{
    uint32_t coop = 0;
    for (uint64_t pending = nleaves; pending > 1; pending = (pending + arity - 1) / arity)
        coop += (pending + arity - 1) / arity <= thr;
    return coop;
}

// The plan job j runs under: `plan` itself, or with NTT fields 0 (a tuned batch)
// the shape's wisdom (the defaults for a shape without an entry).
static Pil2CommitPlan job_plan(const Pil2CommitJob &j, const Pil2CommitPlan &plan)
{
    if (plan.ntt_tile_log) return plan;
    Pil2CommitPlan p = {plan.n_streams, kDefaultPlan.ntt_tile_log, kDefaultPlan.ntt_radix_log,
                        kDefaultPlan.coop_max_parents};
    auto a = g_lde_wisdom.find(lde_key(j));
    if (a != g_lde_wisdom.end()) {
        p.ntt_tile_log = a->second.tile_log;
        p.ntt_radix_log = a->second.radix_log;
    }
    auto b = g_level_wisdom.find(level_key(j));
    if (b != g_level_wisdom.end()) p.coop_max_parents = b->second;
    return p;
}

// Best of two timed runs of f on st (after one untimed run).
static int64_t time_best(const std::function<int64_t()> &f, hipStream_t st, hipEvent_t e0, hipEvent_t e1,
                         float *best)
{
    int64_t rc = f();
    *best = 1e30f;
    for (int k = 0; k < 2 && !rc; k++) {
        HIP_TRY(hipEventRecord(e0, st), "tune event");
        if ((rc = f())) break;
        HIP_TRY(hipEventRecord(e1, st), "tune event");
        HIP_TRY(hipEventSynchronize(e1), "tune sync");
        float ms = 0;
        HIP_TRY(hipEventElapsedTime(&ms, e0, e1), "tune time");
        *best = std::min(*best, ms);
    }
    return rc;
}

// Measures every shape of the batch that has no wisdom yet. The runs write the
// job's own LDE and tree (with the same values the commit then recomputes).
static int64_t tune_shapes(const Pil2CommitJob *jobs, uint64_t n, hipStream_t st, hipEvent_t e0, hipEvent_t e1)
{
    for (uint64_t i = 0; i < n; i++) {
        const Pil2CommitJob &j = jobs[i];
        uint64_t lk = lde_key(j);
        if (!g_lde_wisdom.count(lk)) {
            NttPlan best = {kDefaultPlan.ntt_tile_log, kDefaultPlan.ntt_radix_log};
            float best_ms = 1e30f;
            std::set<uint64_t> seen;
            for (uint32_t tile = 12; tile <= NTT_LOG_TILE; tile++)
                for (uint32_t radix = 2; radix <= 4; radix++) {
                    NttPlan np = {tile, radix};
                    uint64_t sig = ntt_lde_plan_signature((uint32_t)j.n_bits, (uint32_t)j.n_bits_ext, &np);
                    if (!sig || !seen.insert(sig).second) continue;
                    Pil2CommitPlan cp = {1, tile, radix, kDefaultPlan.coop_max_parents};
                    float ms = 0;
                    int64_t rc = time_best(
                        [&] { return lde_impl(j.d_lde, j.d_trace, j.d_scratch, j.n_bits, j.n_bits_ext, j.n_cols, st,
                                              nullptr, cp); },
                        st, e0, e1, &ms);
                    if (rc) return rc;
                    if (ms < best_ms) best_ms = ms, best = np;
                }
            g_lde_wisdom[lk] = best;
        }
        uint64_t vk = level_key(j);
        if (!g_level_wisdom.count(vk)) {
            uint64_t NExt = 1ull << j.n_bits_ext;
            uint32_t best = kDefaultPlan.coop_max_parents;
            float best_ms = 1e30f;
            std::set<uint32_t> seen;
            for (uint32_t thr : kLevelThresholds) {
                if (!seen.insert(level_schedule(NExt, j.arity, thr)).second) continue;
                float ms = 0;
                int64_t rc = time_best(
                    [&]() -> int64_t {
                        int e = p2_launch_tree_levels((uint32_t)(4 * j.arity), j.d_nodes, NExt, thr, st);
                        return e ? fail(-100 - e, "tree level launch failed") : 0;
                    },
                    st, e0, e1, &ms);
                if (rc) return rc;
                if (ms < best_ms) best_ms = ms, best = thr;
            }
            g_level_wisdom[vk] = best;
        }
    }
    return 0;
}

int64_t pil2hip_batch_shape_plan(uint64_t nBits, uint64_t nBitsExt, uint64_t nCols, uint64_t arity,
                                 Pil2CommitPlan *out)
{
    if (!out) return fail(-13, "null plan");
    std::lock_guard<std::mutex> lk(g_batch_mu);
    Pil2CommitJob j = {nullptr, nullptr, nullptr, nullptr, nBits, nBitsExt, nCols, arity};
    bool known = g_lde_wisdom.count(lde_key(j)) && g_level_wisdom.count(level_key(j));
    *out = job_plan(j, Pil2CommitPlan{0, 0, 0, 0});
    return known ? 1 : 0;
}

static bool plan_eq(const Pil2CommitPlan &a, const Pil2CommitPlan &b)
{
    return a.n_streams == b.n_streams && a.ntt_tile_log == b.ntt_tile_log && a.ntt_radix_log == b.ntt_radix_log &&
           a.coop_max_parents == b.coop_max_parents;
}

// Captures `plan` and times one replay (after a warm one) when `timed`.
static int64_t batch_try(const Pil2CommitJob *jobs, uint64_t n, const Pil2CommitPlan &plan, uint64_t *d_roots,
                         hipStream_t st, bool timed, hipEvent_t e0, hipEvent_t e1, hipGraphExec_t *ex, float *ms)
{
    int64_t rc = batch_capture(jobs, n, plan, d_roots, ex);
    if (rc || !timed) return rc;
    hipError_t te = hipGraphLaunch(*ex, st);
    if (te == hipSuccess) te = hipEventRecord(e0, st);
    if (te == hipSuccess) te = hipGraphLaunch(*ex, st);
    if (te == hipSuccess) te = hipEventRecord(e1, st);
    if (te == hipSuccess) te = hipEventSynchronize(e1);
    if (te == hipSuccess) te = hipEventElapsedTime(ms, e0, e1);
    if (te != hipSuccess) {
        (void)hipStreamSynchronize(st);  // a launch may have been queued before the failure
        (void)hipGraphExecDestroy(*ex);
        *ex = nullptr;
        return fail(-100 - (int64_t)te, std::string("autotune: ") + hipGetErrorString(te));
    }
    return 0;
}

int64_t pil2hip_commit_batch_plan(const Pil2CommitJob *jobs, uint64_t n_jobs, const Pil2CommitPlan *plan,
                                  Pil2CommitPlan *chosen, uint64_t *d_roots, void *stream)
{
    int64_t rc = check_family();
    if (rc) return rc;
    if (n_jobs == 0) return 0;
    if (!jobs) return fail(-13, "null jobs");
    Pil2CommitPlan want = plan ? *plan : Pil2CommitPlan{0, 0, 0, 0};
    if (want.n_streams > kMaxBatchStreams) return fail(-27, "n_streams > 16");
    if (want.ntt_tile_log && (want.ntt_tile_log < 12 || want.ntt_tile_log > NTT_LOG_TILE))
        return fail(-27, "ntt_tile_log must be 12..14");
    if (want.ntt_radix_log && (want.ntt_radix_log < 2 || want.ntt_radix_log > 4))
        return fail(-27, "ntt_radix_log must be 2..4");
    for (uint64_t i = 0; i < n_jobs; i++) {
        const Pil2CommitJob &j = jobs[i];
        if ((rc = check_arity(j.arity))) return rc;
        if (j.n_bits_ext < j.n_bits || j.n_bits_ext > NTT_MAX_LOG) return fail(-10, "bad nBits/nBitsExt in a job");
        if (!j.d_trace || !j.d_lde || !j.d_nodes || !j.d_scratch) return fail(-13, "null buffer in a job");
    }
    // all fields 0: tune everything (branch count per job list, NTT tile and radix
    // and level threshold per shape); otherwise plan fields left 0 take the defaults
    bool tune_all = !want.n_streams && !want.ntt_tile_log && !want.ntt_radix_log && !want.coop_max_parents;
    if (!tune_all) {
        if (!want.ntt_tile_log) want.ntt_tile_log = kDefaultPlan.ntt_tile_log;
        if (!want.ntt_radix_log) want.ntt_radix_log = kDefaultPlan.ntt_radix_log;
        if (!want.coop_max_parents) want.coop_max_parents = kDefaultPlan.coop_max_parents;
    }
    NttTablesHost *tb;
    if ((rc = ensure_tables(&tb))) return rc;  // allocations happen before any capture
    std::lock_guard<std::mutex> lk(g_batch_mu);
    if ((rc = batch_resources())) return rc;
    std::vector<uint64_t> key;
    key.push_back((uint64_t)(uintptr_t)d_roots);
    for (uint64_t i = 0; i < n_jobs; i++) {
        const Pil2CommitJob &j = jobs[i];
        key.insert(key.end(), {(uint64_t)(uintptr_t)j.d_trace, (uint64_t)(uintptr_t)j.d_lde,
                               (uint64_t)(uintptr_t)j.d_nodes, (uint64_t)(uintptr_t)j.d_scratch, j.n_bits,
                               j.n_bits_ext, j.n_cols, j.arity});
    }
    hipStream_t st = (hipStream_t)stream;
    auto matches = [&](const Pil2CommitPlan &p) {
        if (tune_all) return p.ntt_tile_log == 0;  // a graph built from shape wisdom
        return (!want.n_streams || p.n_streams == want.n_streams) && p.ntt_tile_log == want.ntt_tile_log &&
               p.ntt_radix_log == want.ntt_radix_log && p.coop_max_parents == want.coop_max_parents;
    };
    for (auto &g : g_batch_graphs)
        if (g.key == key && matches(g.plan)) {
            HIP_TRY(launch_cached(g.exec, g.done, st), "graph launch");
            if (chosen) *chosen = g.plan;
            return g.plan.n_streams;
        }
    // the branch count is tuned when it was 0; NTT fields 0 (tune_all) = per-shape wisdom
    Pil2CommitPlan cur = {want.n_streams ? want.n_streams : (uint32_t)std::min<uint64_t>(4, n_jobs),
                          want.ntt_tile_log, want.ntt_radix_log, want.coop_max_parents};
    std::vector<std::vector<Pil2CommitPlan>> stages;
    if (!want.n_streams) {
        std::vector<Pil2CommitPlan> c;
        for (uint32_t s = 1; s <= kMaxBatchStreams && s <= std::max<uint64_t>(n_jobs, 1); s *= 2)
            c.push_back({s, 0, 0, 0});
        stages.push_back(c);
    }
    hipEvent_t e0, e1;
    HIP_TRY(hipEventCreate(&e0), "event");
    HIP_TRY(hipEventCreate(&e1), "event");
    hipGraphExec_t best = nullptr;
    float best_ms = 1e30f;
    Pil2CommitPlan best_plan = cur;
    if (tune_all) rc = tune_shapes(jobs, n_jobs, st, e0, e1);  // eager runs, before any capture
    bool tuning = !stages.empty();
    if (!tuning && !rc) {
        float ms = 0;
        rc = batch_try(jobs, n_jobs, cur, d_roots, st, false, e0, e1, &best, &ms);
    }
    for (size_t si = 0; si < stages.size() && !rc; si++) {
        Pil2CommitPlan base = best_plan;
        for (const Pil2CommitPlan &delta : stages[si]) {
            Pil2CommitPlan cand = base;
            if (delta.n_streams) cand.n_streams = delta.n_streams;
            if (delta.ntt_tile_log) cand.ntt_tile_log = delta.ntt_tile_log;
            if (delta.ntt_radix_log) cand.ntt_radix_log = delta.ntt_radix_log;
            if (delta.coop_max_parents) cand.coop_max_parents = delta.coop_max_parents;
            if (best && plan_eq(cand, best_plan)) continue;  // already timed
            hipGraphExec_t ex = nullptr;
            float ms = 0;
            if ((rc = batch_try(jobs, n_jobs, cand, d_roots, st, true, e0, e1, &ex, &ms))) break;
            if (ms < best_ms) {
                if (best) (void)hipGraphExecDestroy(best);
                best = ex;
                best_ms = ms;
                best_plan = cand;
            } else {
                (void)hipGraphExecDestroy(ex);
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) {
        if (best) (void)hipGraphExecDestroy(best);
        return rc;
    }
    hipEvent_t done = nullptr;
    hipError_t de = hipEventCreateWithFlags(&done, hipEventDisableTiming);
    if (de != hipSuccess) {
        (void)hipGraphExecDestroy(best);
        return fail(-100 - (int64_t)de, "graph event");
    }
    if (g_batch_graphs.size() >= 8) {  // small cache: drop the oldest graph once its last launch has ended
        retire_exec(g_batch_graphs.front().exec, g_batch_graphs.front().done);
        g_batch_graphs.erase(g_batch_graphs.begin());
    }
    g_batch_graphs.push_back({key, best_plan, best, done});
    HIP_TRY(launch_cached(best, done, st), "graph launch");
    if (chosen) *chosen = best_plan;
    return best_plan.n_streams;
}

int64_t pil2hip_commit_batch(const Pil2CommitJob *jobs, uint64_t n_jobs, uint32_t n_streams, uint64_t *d_roots,
                             void *stream)
{
    // n_streams = 0 tunes the whole plan; otherwise the default plan with n_streams branches
    if (n_streams == 0) return pil2hip_commit_batch_plan(jobs, n_jobs, nullptr, nullptr, d_roots, stream);
    Pil2CommitPlan p = kDefaultPlan;
    p.n_streams = n_streams;
    return pil2hip_commit_batch_plan(jobs, n_jobs, &p, nullptr, d_roots, stream);
}

void pil2hip_batch_release(void)
{
    std::lock_guard<std::mutex> lk(g_batch_mu);
    for (auto &g : g_batch_graphs) retire_exec(g.exec, g.done);
    g_batch_graphs.clear();
}

uint64_t pil2hip_proof_levels(uint64_t nRows, uint64_t arity, uint64_t llv)
{
    if (arity < 2) return 0;
    uint64_t cut = 1;
    for (uint64_t k = 0; k < llv && cut <= nRows; k++) cut *= arity;
    uint64_t n = nRows, levels = 0;
    while (n > cut) {
        n = (n + arity - 1) / arity;
        levels++;
    }
    return levels;
}

uint64_t pil2hip_proof_words(uint64_t nRows, uint64_t nCols, uint64_t arity, uint64_t llv)
{
    return nCols + (arity - 1) * 4 * pil2hip_proof_levels(nRows, arity, llv);
}

int64_t pil2hip_group_proofs(uint64_t *d_out, const uint64_t *d_src, uint64_t rowStride, uint64_t colStride,
                             uint64_t nCols, const uint64_t *d_nodes, uint64_t nRows, uint64_t arity, uint64_t llv,
                             const uint64_t *d_idx, uint64_t nQueries, void *stream)
{
    int64_t rc = check_arity(arity);
    if (rc) return rc;
    if (nQueries == 0) return 0;
    if (!d_out || !d_nodes || !d_idx || (nCols && !d_src)) return fail(-13, "null buffer");
    if (llv > 8) return fail(-28, "last level verification > 8");
    uint64_t cut = 1;
    for (uint64_t k = 0; k < llv && cut <= nRows; k++) cut *= arity;
    uint64_t pw = pil2hip_proof_words(nRows, nCols, arity, llv);
    int e = p2_launch_group_proofs(d_src, rowStride, colStride, nCols, d_nodes, nRows, arity, cut, d_idx, nQueries, pw,
                                   d_out, (hipStream_t)stream);
    if (e) return fail(-100 - e, "group proof launch failed");
    return 0;
}

// Grinding (Poseidon2Goldilocks<8>::grinding, poseidon2_goldilocks.cpp:177-232,
// called by gen_proof through starks_api_internal.hpp:21): the smallest nonce
// whose permuted state word 0 is below 2^(64 - nBits), searched in windows of
// increasing nonces until one holds a valid nonce, up to the reference's bound
// N = ceil(-128 ln 2 / ln(1 - 2^-nBits)).
static uint64_t *g_grind_found[kMaxDevices];
static std::mutex g_grind_mu;

int64_t pil2hip_grinding(uint64_t *nonce, const uint64_t *challenge, uint32_t nBits, void *stream)
{
    if (!nonce || !challenge) return fail(-13, "null argument");
    if (nBits < 1 || nBits > 63) return fail(-36, "grinding bits must be 1..63");
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev), "hipGetDevice");
    if (dev < 0 || dev >= kMaxDevices) return fail(-4, "device index out of range");
    std::lock_guard<std::mutex> lk(g_grind_mu);
    if (!g_grind_found[dev]) HIP_TRY(hipMalloc(&g_grind_found[dev], 8), "hipMalloc(grinding)");
    uint64_t *d_found = g_grind_found[dev];
    const double eps = std::ldexp(1.0, -(int)nBits);
    const double total = -128.0 * std::log(2.0) / std::log1p(-eps);
    const uint64_t limit = total >= 1.8e19 ? UINT64_MAX : (uint64_t)std::ceil(total);
    const uint64_t level = 1ull << (64 - nBits);
    hipStream_t st = (hipStream_t)stream;
    uint64_t none = UINT64_MAX, got = UINT64_MAX;
    HIP_TRY(hipMemcpyAsync(d_found, &none, 8, hipMemcpyHostToDevice, st), "grinding reset");
    const uint64_t window = 1ull << 22;
    for (uint64_t base = 0; base < limit; base += window) {
        uint64_t count = limit - base < window ? limit - base : window;
        int e = p2_launch_grind(challenge, level, base, count, d_found, st);
        if (e) return fail(-100 - e, "grinding launch failed");
        HIP_TRY(hipMemcpyAsync(&got, d_found, 8, hipMemcpyDeviceToHost, st), "grinding result");
        HIP_TRY(hipStreamSynchronize(st), "grinding sync");
        if (got != UINT64_MAX) {
            *nonce = got;
            return 0;
        }
    }
    return fail(-37, "grinding: no valid nonce below the bound");
}

// ------------------------------------------------------------- transcript --
struct TranscriptHip {
    P2TranscriptState *d = nullptr;  // device state
    uint64_t *d_io = nullptr;        // device staging for inputs and outputs
    uint64_t io_words = 0;
    uint32_t width = 0;
};

static int64_t tr_io(TranscriptHip *t, uint64_t n)
{
    if (n <= t->io_words) return 0;
    if (t->d_io) (void)hipFree(t->d_io);
    t->d_io = nullptr;
    t->io_words = 0;
    uint64_t w = n < 64 ? 64 : n;
    HIP_TRY(hipMalloc(&t->d_io, 8 * w), "hipMalloc(transcript io)");
    t->io_words = w;
    return 0;
}

void *pil2hip_transcript_create(uint64_t arity)
{
    if (check_arity(arity)) return nullptr;
    TranscriptHip *t = new TranscriptHip();
    t->width = (uint32_t)(4 * arity);
    P2TranscriptState zero;
    memset(&zero, 0, sizeof(zero));
    if (hipMalloc(&t->d, sizeof(P2TranscriptState)) != hipSuccess ||
        hipMemcpy(t->d, &zero, sizeof(zero), hipMemcpyHostToDevice) != hipSuccess || tr_io(t, 64)) {
        if (t->d) (void)hipFree(t->d);
        delete t;
        fail(-8, "transcript: device allocation failed");
        return nullptr;
    }
    return t;
}

void pil2hip_transcript_destroy(void *handle)
{
    TranscriptHip *t = (TranscriptHip *)handle;
    if (!t) return;
    (void)hipDeviceSynchronize();
    if (t->d) (void)hipFree(t->d);
    if (t->d_io) (void)hipFree(t->d_io);
    delete t;
}

static int64_t tr_run(TranscriptHip *t, uint64_t *io, uint64_t n, int op, hipStream_t st)
{
    int e = p2_launch_transcript(t->width, t->d, io, n, op, st);
    if (e) return fail(-100 - e, "transcript launch failed");
    return 0;
}

int64_t pil2hip_transcript_put(void *handle, const uint64_t *d_in, uint64_t n, void *stream)
{
    TranscriptHip *t = (TranscriptHip *)handle;
    if (!t || (n && !d_in)) return fail(-13, "null argument");
    if (n == 0) return 0;
    return tr_run(t, (uint64_t *)d_in, n, 0, (hipStream_t)stream);
}

int64_t pil2hip_transcript_put_host(void *handle, const uint64_t *h_in, uint64_t n, void *stream)
{
    TranscriptHip *t = (TranscriptHip *)handle;
    if (!t || (n && !h_in)) return fail(-13, "null argument");
    if (n == 0) return 0;
    int64_t rc = tr_io(t, n);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    HIP_TRY(hipMemcpyAsync(t->d_io, h_in, 8 * n, hipMemcpyHostToDevice, st), "transcript input copy");
    if ((rc = tr_run(t, t->d_io, n, 0, st))) return rc;
    HIP_TRY(hipStreamSynchronize(st), "transcript sync");  // the staging buffer is reused
    return 0;
}

static int64_t tr_read(TranscriptHip *t, uint64_t *h_out, uint64_t n, int op, hipStream_t st)
{
    int64_t rc = tr_io(t, n);
    if (rc) return rc;
    if ((rc = tr_run(t, t->d_io, n, op, st))) return rc;
    HIP_TRY(hipMemcpyAsync(h_out, t->d_io, 8 * n, hipMemcpyDeviceToHost, st), "transcript output copy");
    HIP_TRY(hipStreamSynchronize(st), "transcript sync");
    return 0;
}

int64_t pil2hip_transcript_get_field(void *handle, uint64_t *h_out, void *stream)
{
    TranscriptHip *t = (TranscriptHip *)handle;
    if (!t || !h_out) return fail(-13, "null argument");
    return tr_read(t, h_out, 3, 1, (hipStream_t)stream);
}

int64_t pil2hip_transcript_get_state(void *handle, uint64_t *h_out, uint64_t n, void *stream)
{
    TranscriptHip *t = (TranscriptHip *)handle;
    if (!t || !h_out) return fail(-13, "null argument");
    if (n > t->width) return fail(-38, "transcript state has 4*arity words");
    return tr_read(t, h_out, n, 2, (hipStream_t)stream);
}

int64_t pil2hip_transcript_get_permutations(void *handle, uint64_t *h_out, uint64_t n, uint64_t nBits, void *stream)
{
    TranscriptHip *t = (TranscriptHip *)handle;
    if (!t || (n && !h_out)) return fail(-13, "null argument");
    if (n == 0) return 0;
    if (nBits < 1 || nBits > 63) return fail(-39, "permutation bits must be 1..63");
    uint64_t nfields = (n * nBits - 1) / 63 + 1;
    std::vector<uint64_t> f(nfields);
    int64_t rc = tr_read(t, f.data(), nfields, 1, (hipStream_t)stream);
    if (rc) return rc;
    uint64_t cur = 0, bit = 0;  // 63 bits per field, lowest first (transcriptGL.cpp:92-118)
    for (uint64_t i = 0; i < n; i++) {
        uint64_t a = 0;
        for (uint64_t j = 0; j < nBits; j++) {
            if ((f[cur] >> bit) & 1) a |= 1ull << j;
            if (++bit == 63) {
                bit = 0;
                cur++;
            }
        }
        h_out[i] = a;
    }
    return 0;
}

void pil2hip_timing_enable(bool on) { g_timing = on; }

uint64_t pil2hip_timing_read(double *phase_ms, bool reset)
{
    uint64_t n = g_timed_calls;
    for (int i = 0; i < kPhases; i++) {
        if (phase_ms) phase_ms[i] = g_phase_ms[i];
        if (reset) g_phase_ms[i] = 0;
    }
    if (reset) g_timed_calls = 0;
    return n;
}

int64_t pil2hip_get_root(uint64_t *root, const uint64_t *d_nodes, uint64_t nRows, uint64_t arity, void *stream)
{
    uint64_t n = pil2hip_tree_num_elements(nRows, arity);
    if (n < 4) return fail(-15, "empty tree");
    HIP_TRY(hipMemcpyAsync(root, d_nodes + n - 4, 32, hipMemcpyDeviceToHost, (hipStream_t)stream), "root copy");
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream), "root sync");
    return 0;
}

// MerkleTreeGL::getLevel (merkleTreeGL.cpp:84-101): the first level with at most
// arity^llv nodes, zero-padded to arity^llv digests (what FRI::merkelize stores as
// last_levels, fri.hpp:108). llv = 0 leaves h_level untouched, as the reference.
uint64_t pil2hip_level_words(uint64_t arity, uint64_t llv)
{
    if (arity < 2 || llv == 0) return 0;
    uint64_t lim = 1;
    for (uint64_t i = 0; i < llv; i++) {
        if (lim > (UINT64_MAX >> 2) / arity) return 0;
        lim *= arity;
    }
    return lim * 4;
}

int64_t pil2hip_get_level(uint64_t *h_level, const uint64_t *d_nodes, uint64_t nRows, uint64_t arity, uint64_t llv,
                          void *stream)
{
    int64_t rc = check_arity(arity);
    if (rc) return rc;
    if (llv == 0) return 0;
    uint64_t words = pil2hip_level_words(arity, llv);
    if (!words) return fail(-15, "arity^llv out of range");
    if (!h_level || !d_nodes) return fail(-13, "null buffer");
    uint64_t lim = words / 4, n = nRows, offset = 0;
    while (n > lim) {
        n = (n + arity - 1) / arity;
        offset += n * arity * 4;
    }
    hipStream_t st = (hipStream_t)stream;
    if (n) HIP_TRY(hipMemcpyAsync(h_level, d_nodes + offset, n * 4 * 8, hipMemcpyDeviceToHost, st), "level copy");
    HIP_TRY(hipStreamSynchronize(st), "level sync");
    for (uint64_t i = n * 4; i < words; i++) h_level[i] = 0;
    return 0;
}

int64_t pil2hip_fri_fold(uint64_t *d_out, const uint64_t *d_pol, const uint64_t *challenge, uint64_t step,
                         uint64_t nBitsExt, uint64_t prevBits, uint64_t currentBits, void *stream)
{
    if (step == 0) return 0;
    if (!challenge) return fail(-13, "null challenge");
    int e = fri_launch_fold(d_out, d_pol, challenge, nBitsExt, prevBits, currentBits, fri_tables((hipStream_t)stream),
                            (hipStream_t)stream);
    if (e == -1) return fail(-16, "invalid fold bits");
    if (e == -2) return fail(-17, "fold ratio must be 2^1 .. 2^10");
    if (e) return fail(-100 - e, "fold launch failed");
    return 0;
}

int64_t pil2hip_fri_transpose(uint64_t *d_aux, const uint64_t *d_pol, uint64_t degree, uint64_t tBits, void *stream)
{
    if (tBits > 63 || (degree >> tBits) == 0 || ((degree >> tBits) << tBits) != degree)
        return fail(-18, "degree must be a multiple of 2^tBits");
    int e = fri_launch_transpose(d_aux, d_pol, degree, (uint32_t)tBits, (hipStream_t)stream);
    if (e) return fail(-100 - e, "transpose launch failed");
    return 0;
}

// ---------------------------------------------------------- FRI layer sequence --
// The FRI loop of gen_proof: fold step s > 0 (fri.hpp:32-99) into d_work, then,
// below the last step, getTransposed + merkelize of the layer (fri.hpp:102-109).
// Every launch reads its inputs from device memory (the challenges included), so
// the whole sequence is captured once per argument set and replayed: the layers
// are a chain of short, latency-bound kernels, and one graph launch removes the
// host launch gaps between them.
struct FriGraph {
    std::vector<uint64_t> key;
    hipGraphExec_t exec;
    hipEvent_t done;  // recorded after every launch (see launch_cached)
};
static std::vector<FriGraph> g_fri_graphs;
static std::mutex g_fri_mu;
static hipStream_t g_fri_capture = nullptr;
// Graph branches for the layer trees: the challenges are all in device memory
// before the sequence starts, so a layer's tree (leaves and levels, on its own
// source and node buffers) depends only on that layer's transpose, and the next
// fold only on the transpose having read d_work. The trees of different layers
// then run side by side instead of one after another; each is a chain of
// latency-bound levels.
static constexpr uint32_t kFriBranches = 4;
static hipStream_t g_fri_side[kFriBranches];
static std::vector<hipEvent_t> g_fri_fork;         // one per layer
static hipEvent_t g_fri_join[kFriBranches];

static int64_t fri_branch_resources(uint64_t layers)
{
    if (!g_fri_side[0]) {
        for (uint32_t b = 0; b < kFriBranches; b++) {
            HIP_TRY(hipStreamCreateWithFlags(&g_fri_side[b], hipStreamNonBlocking), "FRI branch stream");
            HIP_TRY(hipEventCreateWithFlags(&g_fri_join[b], hipEventDisableTiming), "FRI join event");
        }
    }
    while (g_fri_fork.size() < layers) {
        hipEvent_t e = nullptr;
        HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming), "FRI fork event");
        g_fri_fork.push_back(e);
    }
    return 0;
}

// branches: each layer's tree goes on branch stream (layer mod kFriBranches),
// forked from st after the layer's transpose and joined into st at the end (used
// under graph capture, where the fork and join become graph edges).
static int64_t fri_enqueue(const uint64_t *d_pol, uint64_t *d_work, const uint64_t *d_ch, uint64_t nbe,
                           const uint64_t *steps, uint64_t n, uint64_t arity, uint64_t *const *d_sources,
                           uint64_t *const *d_nodes, hipStream_t st, bool branches)
{
    uint32_t width = (uint32_t)(4 * arity);
    bool used[kFriBranches] = {};
    int64_t rc = 0;
    for (uint64_t s = 0; s < n && !rc; s++) {
        if (s > 0) {
            int e = fri_launch_fold2(d_work, s == 1 ? d_pol : d_work, nullptr, d_ch + 3 * s, nbe, steps[s - 1],
                                     steps[s], fri_tables(st), st);
            if (e) rc = fail(-100 - e, "fold launch failed");
        }
        if (!rc && s + 1 < n) {
            uint64_t rows = 1ull << steps[s + 1], cols = 3ull << (steps[s] - steps[s + 1]);
            int e = fri_launch_transpose(d_sources[s], s == 0 ? d_pol : d_work, 1ull << steps[s],
                                         (uint32_t)steps[s + 1], st);
            hipStream_t ts = st;
            if (!e && branches) {
                uint32_t b = (uint32_t)(s % kFriBranches);
                hipError_t fe = hipEventRecord(g_fri_fork[s], st);
                if (fe == hipSuccess) fe = hipStreamWaitEvent(g_fri_side[b], g_fri_fork[s], 0);
                if (fe != hipSuccess) {
                    rc = fail(-100 - (int64_t)fe, std::string("FRI fork: ") + hipGetErrorString(fe));
                    break;
                }
                ts = g_fri_side[b];
                used[b] = true;
            }
            if (!e) e = p2_launch_leaves(width, d_sources[s], cols, 1, cols, rows, d_nodes[s], ts);
            if (!e) e = p2_launch_tree_levels(width, d_nodes[s], rows, P2_COOP_MAX_PARENTS, ts);
            if (e) rc = fail(-100 - e, "layer commit launch failed");
        }
    }
    for (uint32_t b = 0; b < kFriBranches; b++) {  // join every forked branch, also after an error
        if (!used[b]) continue;
        hipError_t je = hipEventRecord(g_fri_join[b], g_fri_side[b]);
        if (je == hipSuccess) je = hipStreamWaitEvent(st, g_fri_join[b], 0);
        if (je != hipSuccess && !rc) rc = fail(-100 - (int64_t)je, std::string("FRI join: ") + hipGetErrorString(je));
    }
    return rc;
}

int64_t pil2hip_fri_layers(const uint64_t *d_pol, uint64_t *d_work, const uint64_t *d_challenges, uint64_t nBitsExt,
                           const uint64_t *steps, uint64_t nSteps, uint64_t arity, uint64_t *const *d_sources,
                           uint64_t *const *d_nodes, bool graph, void *stream)
{
    int64_t rc = check_family();
    if (rc) return rc;
    if ((rc = check_arity(arity))) return rc;
    if (!steps || nSteps == 0) return fail(-16, "no FRI steps");
    if (steps[0] != nBitsExt || nBitsExt > 32) return fail(-16, "steps[0] must equal nBitsExt (<= 32)");
    for (uint64_t s = 1; s < nSteps; s++)
        if (steps[s] >= steps[s - 1] || steps[s - 1] - steps[s] > 10)
            return fail(-17, "FRI steps must decrease by 1 .. 10 bits");
    if (!d_pol || (nSteps > 1 && (!d_work || !d_challenges))) return fail(-13, "null buffer");
    if (nSteps > 1 && (!d_sources || !d_nodes)) return fail(-13, "null layer buffers");
    for (uint64_t s = 0; s + 1 < nSteps; s++)
        if (!d_sources[s] || !d_nodes[s]) return fail(-13, "null layer buffer");
    hipStream_t st = (hipStream_t)stream;
    (void)fri_tables(st);  // set up before the capture below, which only looks them up
    if (!graph)
        return fri_enqueue(d_pol, d_work, d_challenges, nBitsExt, steps, nSteps, arity, d_sources, d_nodes, st, false);
    std::vector<uint64_t> key = {(uint64_t)(uintptr_t)d_pol, (uint64_t)(uintptr_t)d_work,
                                 (uint64_t)(uintptr_t)d_challenges, nBitsExt, arity};
    for (uint64_t s = 0; s < nSteps; s++) key.push_back(steps[s]);
    for (uint64_t s = 0; s + 1 < nSteps; s++) {
        key.push_back((uint64_t)(uintptr_t)d_sources[s]);
        key.push_back((uint64_t)(uintptr_t)d_nodes[s]);
    }
    std::lock_guard<std::mutex> lk(g_fri_mu);
    for (auto &g : g_fri_graphs)
        if (g.key == key) {
            HIP_TRY(launch_cached(g.exec, g.done, st), "FRI graph launch");
            return 0;
        }
    if (!g_fri_capture) HIP_TRY(hipStreamCreateWithFlags(&g_fri_capture, hipStreamNonBlocking), "FRI capture stream");
    if ((rc = fri_branch_resources(nSteps))) return rc;
    HIP_TRY(hipStreamBeginCapture(g_fri_capture, hipStreamCaptureModeThreadLocal), "FRI begin capture");
    t_capturing = true;
    rc = fri_enqueue(d_pol, d_work, d_challenges, nBitsExt, steps, nSteps, arity, d_sources, d_nodes, g_fri_capture,
                     true);
    t_capturing = false;
    hipGraph_t g = nullptr;
    hipError_t ee = hipStreamEndCapture(g_fri_capture, &g);
    if (rc || ee != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : fail(-100 - (int64_t)ee, std::string("FRI end capture: ") + hipGetErrorString(ee));
    }
    hipGraphExec_t ex = nullptr;
    ee = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ee != hipSuccess) return fail(-100 - (int64_t)ee, std::string("FRI graph instantiate: ") + hipGetErrorString(ee));
    hipEvent_t done = nullptr;
    ee = hipEventCreateWithFlags(&done, hipEventDisableTiming);
    if (ee != hipSuccess) {
        (void)hipGraphExecDestroy(ex);
        return fail(-100 - (int64_t)ee, "FRI graph event");
    }
    if (g_fri_graphs.size() >= 8) {  // small cache: drop the oldest graph once its last launch has ended
        retire_exec(g_fri_graphs.front().exec, g_fri_graphs.front().done);
        g_fri_graphs.erase(g_fri_graphs.begin());
    }
    g_fri_graphs.push_back({key, ex, done});
    HIP_TRY(launch_cached(ex, done, st), "FRI graph launch");
    return 0;
}

void pil2hip_fri_release(void)
{
    std::lock_guard<std::mutex> lk(g_fri_mu);
    for (auto &g : g_fri_graphs) retire_exec(g.exec, g.done);
    g_fri_graphs.clear();
}

int64_t pil2hip_transpose_cm2rm(uint64_t *d_out, const uint64_t *d_in, uint64_t rows, uint64_t cols,
                                uint64_t inColStride, uint64_t outRowStride, void *stream)
{
    if (rows == 0 || cols == 0) return 0;
    if (cols > 65535u * 32u) return fail(-12, "too many columns");
    int e = ntt_launch_cm2rm(d_out, d_in, rows, cols, inColStride, outRowStride, (hipStream_t)stream);
    if (e) return fail(-100 - e, "transpose launch failed");
    return 0;
}

int64_t pil2hip_p2_permute(uint64_t *d_states, uint64_t count, uint32_t width, void *stream)
{
    if (width != 8 && width != 12 && width != 16) return fail(-19, "width must be 8, 12 or 16");
    int e = p2_launch_permute(width, d_states, count, (hipStream_t)stream);
    if (e) return fail(-100 - e, "permute launch failed");
    return 0;
}

int64_t pil2hip_mul_peak(uint64_t *d_out, const uint64_t *d_in, uint64_t lanes, uint32_t iters, void *stream)
{
    if (lanes > (uint64_t)UINT32_MAX * 256) return fail(-12, "too many lanes");
    if (lanes && (!d_out || !d_in)) return fail(-13, "null buffer");
    int e = p2_launch_mul_peak(d_in, d_out, lanes, iters, (hipStream_t)stream);
    if (e) return fail(-100 - e, "product calibration launch failed");
    return 0;
}

int64_t pil2hip_gl_binop(uint64_t *d_out, const uint64_t *d_a, const uint64_t *d_b, uint64_t n, int32_t op,
                         void *stream)
{
    if (op < 0 || op > 3) return fail(-26, "op must be 0..3");
    int e = fri_launch_gl_binop(d_out, d_a, d_b, n, op, (hipStream_t)stream);
    if (e) return fail(-100 - e, "field op launch failed");
    return 0;
}

int64_t pil2hip_fill_uniform(uint64_t *d_out, uint64_t n, uint64_t seed, uint64_t start, void *stream)
{
    int e = ntt_launch_fill_uniform(d_out, n, seed, start, (hipStream_t)stream);
    if (e) return fail(-100 - e, "fill launch failed");
    return 0;
}

// ------------------------------------------------------ write_custom_commit --
static void legacy_error(const char *where)
{
    if (g_abort_on_error) {
        fprintf(stderr, "%s: %s\n", where, t_err.c_str());
        abort();
    }
}

static int64_t write_file(const char *path, const void *root, const uint64_t *trace, uint64_t traceN,
                          const uint64_t *lde_rm, uint64_t ldeN, const uint64_t *nodes, uint64_t nodesN)
{
    FILE *f = fopen(path, "wb");
    if (!f) return fail(-20, std::string("cannot open ") + path);
    bool ok = fwrite(root, 1, 32, f) == 32 && fwrite(trace, 8, traceN, f) == traceN &&
              fwrite(lde_rm, 8, ldeN, f) == ldeN && fwrite(nodes, 8, nodesN, f) == nodesN;
    ok = (fclose(f) == 0) && ok;
    return ok ? 0 : fail(-21, std::string("short write to ") + path);
}

// A host trace committed on the device, with the outputs the CPU backend leaves
// in host memory: the root, and optionally the row-major LDE (NExt x nCols) and
// the Merkle nodes. Synchronous on the d_buffers stream (or the null stream).
static int64_t host_commit(void *root, uint64_t arity, uint64_t nBits, uint64_t nBitsExt, uint64_t nCols,
                           void *d_buffers, const void *buffer, uint64_t *h_lde_rm, uint64_t *h_nodes,
                           const char *who)
{
    int64_t rc = check_family();
    if (rc) return rc;
    if ((rc = check_arity(arity))) return rc;
    if (nBitsExt < nBits) return fail(-10, "nBitsExt < nBits");
    if (nBitsExt > NTT_MAX_LOG) return fail(-11, "nBitsExt > 28 is not supported");
    if (!root || (!buffer && nCols)) return fail(-13, "null buffer");
    uint64_t N = 1ull << nBits, NExt = 1ull << nBitsExt;
    uint64_t nTree = pil2hip_tree_num_elements(NExt, arity);
    hipStream_t st = d_buffers ? ((DeviceBuffersHip *)d_buffers)->stream : (hipStream_t)0;
    uint64_t *d_trace = nullptr, *d_scratch = nullptr, *d_lde = nullptr, *d_nodes = nullptr, *d_rm = nullptr;
    auto cleanup = [&]() {
        (void)hipFree(d_trace); (void)hipFree(d_scratch); (void)hipFree(d_lde); (void)hipFree(d_nodes); (void)hipFree(d_rm);
    };
    uint64_t tsz = N * nCols, lsz = NExt * nCols;
    bool want_rm = h_lde_rm && lsz;
    if (hipMalloc(&d_trace, (tsz ? tsz : 1) * 8) != hipSuccess || hipMalloc(&d_scratch, (tsz ? tsz : 1) * 8) != hipSuccess ||
        hipMalloc(&d_lde, (lsz ? lsz : 1) * 8) != hipSuccess || hipMalloc(&d_nodes, nTree * 8) != hipSuccess ||
        (want_rm && hipMalloc(&d_rm, lsz * 8) != hipSuccess)) {
        cleanup();
        return fail(-22, std::string(who) + ": device allocation failed");
    }
    if (tsz && hipMemcpyAsync(d_trace, buffer, tsz * 8, hipMemcpyHostToDevice, st) != hipSuccess) {
        (void)hipStreamSynchronize(st);
        cleanup();
        return fail(-23, std::string(who) + ": H2D copy failed");
    }
    rc = pil2hip_lde(d_lde, d_trace, d_scratch, nBits, nBitsExt, nCols, st);
    if (!rc) rc = pil2hip_merkelize(d_nodes, d_lde, NExt, nCols, 1, NExt, arity, st);
    if (!rc) rc = pil2hip_get_root((uint64_t *)root, d_nodes, NExt, arity, st);
    if (!rc && want_rm) {
        rc = pil2hip_transpose_cm2rm(d_rm, d_lde, NExt, nCols, NExt, nCols, st);
        if (!rc && hipMemcpyAsync(h_lde_rm, d_rm, lsz * 8, hipMemcpyDeviceToHost, st) != hipSuccess)
            rc = fail(-24, std::string(who) + ": D2H copy failed");
    }
    if (!rc && h_nodes && hipMemcpyAsync(h_nodes, d_nodes, nTree * 8, hipMemcpyDeviceToHost, st) != hipSuccess)
        rc = fail(-24, std::string(who) + ": D2H copy failed");
    if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = fail(-25, "stream sync failed");
    cleanup();  // also after an error: the stream has drained
    return rc;
}

static int64_t write_custom_commit_impl(void *root, uint64_t arity, uint64_t nBits, uint64_t nBitsExt, uint64_t nCols,
                                        void *d_buffers, void *buffer, const char *bufferFile)
{
    bool want_file = bufferFile && bufferFile[0];
    if (!want_file)
        return host_commit(root, arity, nBits, nBitsExt, nCols, d_buffers, buffer, nullptr, nullptr,
                           "write_custom_commit");
    if ((nBitsExt < nBits) || nBitsExt > NTT_MAX_LOG || arity < 2 || arity > 4)  // sizes below must be sane
        return host_commit(root, arity, nBits, nBitsExt, nCols, d_buffers, buffer, nullptr, nullptr,
                           "write_custom_commit");
    uint64_t N = 1ull << nBits, NExt = 1ull << nBitsExt;
    std::vector<uint64_t> lde(NExt * nCols), nodes(pil2hip_tree_num_elements(NExt, arity));
    int64_t rc = host_commit(root, arity, nBits, nBitsExt, nCols, d_buffers, buffer, lde.data(), nodes.data(),
                             "write_custom_commit");
    if (!rc)
        rc = write_file(bufferFile, root, (const uint64_t *)buffer, N * nCols, lde.data(), lde.size(), nodes.data(),
                        nodes.size());
    return rc;
}

int64_t pil2hip_commit_witness(uint64_t *root, uint64_t arity, uint64_t nBits, uint64_t nBitsExt, uint64_t nCols,
                               const uint64_t *trace, uint64_t *aux_trace, uint64_t offset_dst, uint64_t offset_mt,
                               void *d_buffers)
{
    if (!aux_trace) return fail(-13, "null aux_trace");
    if (nBitsExt < nBits || nBitsExt > NTT_MAX_LOG || arity < 2 || arity > 4)
        return host_commit(root, arity, nBits, nBitsExt, nCols, d_buffers, trace, nullptr, nullptr, "commit_witness");
    uint64_t NExt = 1ull << nBitsExt, lsz = NExt * nCols, nTree = pil2hip_tree_num_elements(NExt, arity);
    // the two sections must not overlap (the reference lays them out apart in aux_trace)
    if (lsz && offset_dst < offset_mt + nTree && offset_mt < offset_dst + lsz)
        return fail(-35, "commit_witness: the LDE and Merkle sections of aux_trace overlap");
    return host_commit(root, arity, nBits, nBitsExt, nCols, d_buffers, trace, aux_trace + offset_dst,
                       aux_trace + offset_mt, "commit_witness");
}

void write_custom_commit(void *root, uint64_t arity, uint64_t nBits, uint64_t nBitsExt, uint64_t nCols, void *d_buffers,
                         void *buffer, char *bufferFile)
{
    t_err.clear();
    if (write_custom_commit_impl(root, arity, nBits, nBitsExt, nCols, d_buffers, buffer, bufferFile))
        legacy_error("write_custom_commit");
}

// ---------------------------------------------------------------- const tree --
// The GL const tree buffer (const_pols.hpp:25-44): the LDE of the constant
// polynomials, row-major NExt x nConstants, followed by the Merkle nodes.
uint64_t pil2hip_const_tree_size(uint64_t nBitsExt, uint64_t nConstants, uint64_t arity)
{
    uint64_t NExt = 1ull << nBitsExt;
    return NExt * nConstants + pil2hip_tree_num_elements(NExt, arity);
}

int64_t pil2hip_const_tree(uint64_t *d_tree, const uint64_t *d_constPols, uint64_t *d_scratch, uint64_t nBits,
                           uint64_t nBitsExt, uint64_t nConstants, uint64_t arity, void *stream)
{
    int64_t rc = check_family();
    if (rc) return rc;
    if ((rc = check_arity(arity))) return rc;
    if (nBitsExt < nBits) return fail(-10, "nBitsExt < nBits");
    if (nBitsExt > NTT_MAX_LOG) return fail(-11, "nBitsExt > 28 is not supported");
    if (!d_tree || (nConstants && (!d_constPols || !d_scratch))) return fail(-13, "null buffer");
    uint64_t NExt = 1ull << nBitsExt, C = nConstants;
    hipStream_t st = (hipStream_t)stream;
    uint64_t *lde_cm = d_scratch, *tmp = d_scratch + NExt * C;
    if (C && (rc = lde_impl(lde_cm, d_constPols, tmp, nBits, nBitsExt, C, st, nullptr))) return rc;
    if ((rc = pil2hip_merkelize(d_tree + NExt * C, lde_cm, NExt, C, 1, NExt, arity, stream))) return rc;
    if (C && (rc = pil2hip_transpose_cm2rm(d_tree, lde_cm, NExt, C, NExt, C, stream))) return rc;
    return 0;
}

int64_t pil2hip_calculate_const_tree(void *pConstPols, void *pConstTree, uint64_t nBits, uint64_t nBitsExt,
                                     uint64_t nConstants, uint64_t arity, void *d_buffers)
{
    int64_t rc = check_family();
    if (rc) return rc;
    if ((rc = check_arity(arity))) return rc;
    if (nBitsExt < nBits || nBitsExt > NTT_MAX_LOG) return fail(-11, "unsupported nBits / nBitsExt");
    if (!pConstTree || (nConstants && !pConstPols)) return fail(-13, "null buffer");
    uint64_t N = 1ull << nBits, NExt = 1ull << nBitsExt, C = nConstants;
    uint64_t tsz = pil2hip_const_tree_size(nBitsExt, C, arity), ssz = (N + NExt) * C;
    hipStream_t st = d_buffers ? ((DeviceBuffersHip *)d_buffers)->stream : (hipStream_t)0;
    uint64_t *d_pols = nullptr, *d_tree = nullptr, *d_scratch = nullptr;
    auto cleanup = [&]() { (void)hipFree(d_pols); (void)hipFree(d_tree); (void)hipFree(d_scratch); };
    if (hipMalloc(&d_pols, (N * C ? N * C : 1) * 8) != hipSuccess || hipMalloc(&d_tree, tsz * 8) != hipSuccess ||
        hipMalloc(&d_scratch, (ssz ? ssz : 1) * 8) != hipSuccess) {
        cleanup();
        return fail(-22, "calculate_const_tree: device allocation failed");
    }
    if (N * C && hipMemcpyAsync(d_pols, pConstPols, N * C * 8, hipMemcpyHostToDevice, st) != hipSuccess)
        rc = fail(-23, "calculate_const_tree: H2D copy failed");
    if (!rc) rc = pil2hip_const_tree(d_tree, d_pols, d_scratch, nBits, nBitsExt, C, arity, st);
    if (!rc && hipMemcpyAsync(pConstTree, d_tree, tsz * 8, hipMemcpyDeviceToHost, st) != hipSuccess)
        rc = fail(-24, "calculate_const_tree: D2H copy failed");
    if (hipStreamSynchronize(st) != hipSuccess && !rc) rc = fail(-25, "stream sync failed");
    cleanup();
    return rc;
}

// MerkleTreeGL::writeFile (merkleTreeGL.cpp:338-345): the source then the nodes,
// i.e. the whole const tree buffer in order.
int64_t pil2hip_write_const_tree(const void *pConstTree, uint64_t nBitsExt, uint64_t nConstants, uint64_t arity,
                                 const char *treeFilename)
{
    if (!pConstTree || !treeFilename) return fail(-13, "null buffer");
    uint64_t n = pil2hip_const_tree_size(nBitsExt, nConstants, arity);
    FILE *f = fopen(treeFilename, "wb");
    if (!f) return fail(-20, std::string("cannot open ") + treeFilename);
    bool ok = fwrite(pConstTree, 8, n, f) == n;
    ok = (fclose(f) == 0) && ok;
    return ok ? 0 : fail(-21, std::string("short write to ") + treeFilename);
}

// ConstTree::loadConstTree (const_pols.hpp:66-103), GL branch: read constTreeSize
// words, then compare the root (the last 4 words) with the verification key, a
// JSON array of 4 integers. 1 = loaded and verified, 0 = not, < 0 = a
// verification key that is not 4 integers.
int64_t pil2hip_load_const_tree(void *pConstTree, const char *treeFilename, uint64_t constTreeSize,
                                const char *verkeyFilename)
{
    if (!pConstTree || !treeFilename || constTreeSize < 4) return fail(-13, "null buffer or size below one root");
    FILE *f = fopen(treeFilename, "rb");
    if (!f) return 0;
    bool ok = fread(pConstTree, 8, constTreeSize, f) == constTreeSize;
    fclose(f);
    if (!ok || !verkeyFilename || !verkeyFilename[0]) return ok ? 1 : 0;
    FILE *k = fopen(verkeyFilename, "rb");
    if (!k) return 0;
    std::string text;
    char buf[256];
    size_t got;
    while ((got = fread(buf, 1, sizeof(buf), k)) > 0) text.append(buf, got);
    fclose(k);
    uint64_t vk[4];
    int nv = 0;
    const char *p = text.c_str();
    while (*p && nv < 4) {
        if (*p >= '0' && *p <= '9') {
            char *end;
            vk[nv++] = strtoull(p, &end, 10);
            p = end;
        } else {
            p++;
        }
    }
    if (nv != 4) return fail(-26, std::string("verification key is not 4 integers: ") + verkeyFilename);
    const uint64_t *root = (const uint64_t *)pConstTree + constTreeSize - 4;
    for (int i = 0; i < 4; i++)
        if (root[i] % GL_P != vk[i]) return 0;
    return 1;
}

}  // extern "C"
