// mr3_driver.hip -- host side of the MI355X MRRR solver: the batched device pipeline
// and the drop-in mrrr() entry point (HPAC/mr3smp SRC/mrrr.c:87-296).
//
// A solve runs the kernels of mr3_kernels.hip on one stream (with side streams for the
// singletons of each tree level).  Every launch is sized on the device: the kernels
// read their item counts (blocks, clusters, eigen entries, singletons, wave tasks) from
// device memory and walk them with grid-stride loops, and the block lists are built by
// a kernel, so a grid is only a hint.  The representation tree is walked level by
// level; the host waits for nothing but each level's counters (an event on a small
// copy into pinned memory), which give the size of the level's child representation
// slots, while the side streams go on.  A solve with a fixed number of levels and one
// pool for all slots needs no count at all: that is what a captured graph replays
// (mr3_graph_*).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mr3hip.h"
#include "mr3_internal.h"

#include "mr3_kernels.hip"
#include "mr3_small.hip"

namespace mr3 {

#define MR3_CHECK(x)                                                                    \
    do {                                                                                \
        hipError_t e_ = (x);                                                            \
        if (e_ != hipSuccess) {                                                         \
            if (getenv("MR3_DEBUG"))                                                    \
                fprintf(stderr, "mr3: %s failed: %s (%s:%d)\n", #x, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                            \
            return 2;                                                                   \
        }                                                                               \
    } while (0)

// A device failure of the solve (code 2), reported with its place under MR3_DEBUG.
// Host timestamps of the eager tree walk (a profiling build defines the hook); none by
// default.
#ifndef MR3_HOST_MARK
#define MR3_HOST_MARK(tag, level) ((void)0)
#endif
#define MR3_FAIL()                                                                      \
    do {                                                                                \
        if (getenv("MR3_DEBUG")) fprintf(stderr, "mr3: failure at %s:%d\n", __FILE__, __LINE__); \
        return 2;                                                                       \
    } while (0)

static inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

// Launch-shape parameters of a solve (mr3_set_tuning).  The results do not depend on
// them; the defaults were measured on the batched workload and an autotuner
// (mr3smp_amd/tune.py) can search them for another workload or device.
//   side_streams    singleton side streams (1 .. 3; 0: three with 8 or more hardware
//                   queues, else two)
//   tail_singletons the tail kernel takes over an iteration with at most this many
//                   singletons going on
//   group_lanes     lane budget of the group (multisection) cluster refinements
//   rrf_group_max   levels with more clusters run dlarrf with one lane per cluster
//   singleton_waves most waves of one singleton launch (also bounds its scratch)
//   small_first     small-route matrices of at least this order start with the solve;
//                   the smaller ones wait until the pipeline's root eigenvalues are
//                   refined (0: all start at once)
static std::mutex g_tune_mu;
static mr3_tuning g_tune = {-1, 8192, 262144, 1 << 16, 8192, 0};
static mr3_tuning tuning_snapshot() {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    if (g_tune.side_streams < 0) {   // first use: the environment's overrides
        g_tune.side_streams = 0;
        if (const char* v = getenv("MR3_SIDE_STREAMS")) g_tune.side_streams = std::max(1, std::min(3, atoi(v)));
        if (const char* v = getenv("MR3_SWAVES")) g_tune.singleton_waves = std::max(64, atoi(v));
    }
    return g_tune;
}


// ---- profiling: HIP events around every launch (on the launch stream) and the
// on-device step counters, accumulated over the solves run while enabled.
struct ProfEntry {
    double ms = 0.0;
    long long launches = 0;
};
struct Profiler {
    std::mutex mu;
    std::atomic<int> on{0};
    std::map<std::string, ProfEntry> acc;
    unsigned long long steps[kStepKinds] = {};
};
static Profiler g_prof;

// Timing events are reused: creating two per launch costs the host more than the
// short launches of a level take on the device.
struct EventPool {
    std::vector<hipEvent_t> free;
    hipEvent_t get() {
        if (!free.empty()) {
            hipEvent_t e = free.back();
            free.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }
    void put(hipEvent_t e) {
        if (e) free.push_back(e);
    }
};
static EventPool& event_pool() {
    thread_local EventPool p;
    return p;
}

struct CallEvents {
    bool on = false;
    std::vector<std::pair<const char*, std::pair<hipEvent_t, hipEvent_t>>> ev;
    void collect() {
        if (!on) return;
        std::lock_guard<std::mutex> lk(g_prof.mu);
        for (auto& e : ev) {
            float ms = 0.f;
            if (hipEventElapsedTime(&ms, e.second.first, e.second.second) == hipSuccess) {
                ProfEntry& pe = g_prof.acc[e.first];
                pe.ms += ms;
                pe.launches += 1;
            }
            event_pool().put(e.second.first);
            event_pool().put(e.second.second);
        }
        ev.clear();
        if (g_prof.on.load() != 1) return;   // timing only: the work counters are off
        unsigned long long st[kStepKinds];
        if (hipMemcpyFromSymbol(st, HIP_SYMBOL(g_steps), sizeof(st)) == hipSuccess) {
            for (int i = 0; i < kStepKinds; ++i) g_prof.steps[i] += st[i];
            unsigned long long z[kStepKinds] = {};
            (void)hipMemcpyToSymbol(HIP_SYMBOL(g_steps), z, sizeof(z));
        }
    }
};

#define LAUNCH_ON(STRM, KERN, GRID, BLOCK, ...)                                     \
    do {                                                                            \
        hipEvent_t e0_ = nullptr, e1_ = nullptr;                                    \
        if (ce.on) {                                                                \
            e0_ = event_pool().get();                                               \
            e1_ = event_pool().get();                                               \
            (void)hipEventRecord(e0_, STRM);                                        \
        }                                                                           \
        hipLaunchKernelGGL(KERN, GRID, BLOCK, 0, STRM, __VA_ARGS__);                 \
        if (ce.on) {                                                                \
            (void)hipEventRecord(e1_, STRM);                                        \
            ce.ev.push_back({#KERN, {e0_, e1_}});                                   \
        }                                                                           \
    } while (0)
#define LAUNCH(KERN, GRID, BLOCK, ...) LAUNCH_ON(st, KERN, GRID, BLOCK, __VA_ARGS__)

// The singleton iterations of a tree level run on a side stream of the calling
// thread, beside the cluster work of the same level (they touch disjoint eigenvalue
// slots, Z columns and representations, as the reference's S- and C-tasks do).
// Side streams take consecutive levels in turn, so that the last few singletons of
// level L (the ones that need all ten iterations) run beside the bulk of the next
// levels.  How many pay depends on the hardware queues of the process: HIP gives it
// GPU_MAX_HW_QUEUES of them (4 by default) and streams beyond that share one, so
// that a level can wait for the tail of an unrelated one.  Measured on the batched
// workload: two side streams with 4 queues (three were slower), three with 8 or more
// queues.  The side_streams tuning parameter (or MR3_SIDE_STREAMS) overrides it.
constexpr int kSide = 3;   // most side streams of a solve
constexpr int kTailBigWG = 512;   // workgroups of the large-block tail kernel
static int side_count(int tuned) {
    if (tuned >= 1) return std::min(tuned, kSide);
    if (const char* q = getenv("GPU_MAX_HW_QUEUES"))
        if (atoi(q) >= 8) return 3;
    return 2;
}
struct SideStreams {
    hipStream_t s[kSide] = {};
    bool tried[kSide] = {};
};
static SideStreams& side_streams() {
    thread_local SideStreams ss;
    return ss;
}
static hipStream_t side_stream(int k) {
    SideStreams& ss = side_streams();
    if (!ss.tried[k]) {
        ss.tried[k] = true;
        if (hipStreamCreateWithFlags(&ss.s[k], hipStreamNonBlocking) != hipSuccess) ss.s[k] = nullptr;
    }
    return ss.s[k];
}

// Destroys the calling thread's side streams (a thread that solves and then ends
// calls it, e.g. a server worker; the streams are created again if it solves later).
void release_thread_streams() {
    SideStreams& ss = side_streams();
    for (int k = 0; k < kSide; ++k) {
        if (ss.s[k]) {
            (void)hipStreamSynchronize(ss.s[k]);
            (void)hipStreamDestroy(ss.s[k]);
        }
        ss.s[k] = nullptr;
        ss.tried[k] = false;
    }
}

// Device memory of a solve: one block carved by a bump pointer.  Blocks are kept in a
// process-wide free list between solves, so that a solve does not allocate and
// release its sixty-odd buffers one by one (returning them to the stream-ordered pool
// took about 12 ms per batch of 4096 matrices).  A solve that needs more than its
// block takes the rest from the stream-ordered pool and, at its end, replaces the
// block by one large enough for next time.  Concurrent solves (host threads, server
// batches) each hold their own block.  The list keeps at most 60% of the device's
// memory: a block handed back beyond that is released (a server's workers each grow a
// block to the largest solve they met).
struct Workspace {
    char* base = nullptr;
    size_t cap = 0;
};
static std::mutex g_ws_mu;
static std::vector<Workspace> g_ws_free;
static size_t ws_cache_bytes() {
    static size_t cap = 0;
    static std::once_flag once;
    std::call_once(once, [] {
        size_t fr = 0, tot = 0;
        cap = (hipMemGetInfo(&fr, &tot) == hipSuccess) ? tot / 10 * 6 : ((size_t)64 << 30);
    });
    return cap;
}

struct Alloc {
    hipStream_t st;
    Workspace ws;
    size_t off = 0, need = 0;
    std::vector<void*> extra;   // buffers beyond the block
    bool ok = true;
    bool borrowed = false;      // a block owned by a captured graph: no growth, no release
    Alloc(hipStream_t s, char* base, size_t cap) : st(s), borrowed(true) {
        ws.base = base;
        ws.cap = cap;
    }
    explicit Alloc(hipStream_t s) : st(s) {
        std::lock_guard<std::mutex> lk(g_ws_mu);
        if (!g_ws_free.empty()) {
            // the largest block on the list
            size_t k = 0;
            for (size_t j = 1; j < g_ws_free.size(); ++j)
                if (g_ws_free[j].cap > g_ws_free[k].cap) k = j;
            ws = g_ws_free[k];
            g_ws_free.erase(g_ws_free.begin() + (long)k);
        }
    }
    template <class T>
    T* get(size_t count) {
        if (count == 0) count = 1;
        const size_t bytes = (count * sizeof(T) + 255) & ~(size_t)255;
        need += bytes;
        if (ws.base && off + bytes <= ws.cap) {
            void* p = ws.base + off;
            off += bytes;
            return (T*)p;
        }
        void* p = nullptr;
        if (borrowed || hipMallocAsync(&p, bytes, st) != hipSuccess) { ok = false; return nullptr; }
        extra.push_back(p);
        return (T*)p;
    }
    // An array that the sweeps read: kSweepPad spare doubles on both sides, so that a
    // sweep may read a few entries past either end of a block without an index clamp
    // (the values read there are never used).
    double* get_swept(size_t count) {
        double* p = get<double>(count + 2 * (size_t)kSweepPad);
        return p ? p + kSweepPad : nullptr;
    }
    ~Alloc() {
        if (borrowed) return;
        // every stream of the solve is idle once the main stream is: the side streams
        // were synchronised when the level loop's scope ended
        (void)hipStreamSynchronize(st);
        if (!extra.empty()) {
            for (void* p : extra) (void)hipFreeAsync(p, st);
            if (ws.base) (void)hipFreeAsync(ws.base, st);
            ws = Workspace{};
            const size_t cap = need + need / 4;
            void* p = nullptr;
            if (hipMallocAsync(&p, cap, st) == hipSuccess) {
                ws.base = (char*)p;
                ws.cap = cap;
            }
            // the block is used on other streams by later solves
            (void)hipStreamSynchronize(st);
        }
        if (ws.base) {
            std::lock_guard<std::mutex> lk(g_ws_mu);
            size_t held = 0;
            for (const Workspace& w : g_ws_free) held += w.cap;
            if (held + ws.cap <= ws_cache_bytes()) {
                g_ws_free.push_back(ws);
            } else {
                (void)hipFreeAsync(ws.base, st);
            }
        }
    }
};

static void configure_pool_once() {
    static std::once_flag once;
    std::call_once(once, [] {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return;
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
            uint64_t thr = ~0ULL;   // keep freed memory cached in the pool
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &thr);
        }
    });
}

// Scratch budget of the eigenvector kernel (per batch): a quarter of free memory,
// at most 32 GiB.
// Device memory for the singleton kernels' per-wave checkpoints: a quarter of the free
// memory, at most 32 GiB.  Within it a singleton launch gets at most the tuned number of
// waves, and beyond 2 GiB of checkpoints per side stream no more than the device holds
// at once (a wave takes tasks until none are left, so a wave that can only start after
// others have ended finds none).  Without that bound a server's concurrent solves of
// blocks of 16384 rows each took about 26 GiB and ran the device out of memory.
static size_t singleton_scratch_budget() {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return (size_t)1 << 30;
    size_t b = fr / 4;
    const size_t cap = (size_t)32 << 30;
    return std::min(b, cap);
}

static int singleton_waves_resident() {
    static std::once_flag once;
    static int waves = 0;
    std::call_once(once, [] {
        int dev = 0, per_cu = 0;
        hipDeviceProp_t pr;
        if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&pr, dev) != hipSuccess) return;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_s_iter, 64, 0) != hipSuccess)
            return;
        waves = per_cu * pr.multiProcessorCount;
    });
    return waves > 0 ? waves : 2048;
}

// Pinned host copies of the counters of every tree level of a solve, reused across
// solves (allocating pinned memory per solve would cost more than the solve of a small
// matrix).
constexpr int kMaxLevelsPinned = 66;   // levels 0 .. 64 and the block count
static std::mutex g_pin_mu;
static std::vector<LevelCounters*> g_pin_free;
static LevelCounters* pinned_levels_get() {
    {
        std::lock_guard<std::mutex> lk(g_pin_mu);
        if (!g_pin_free.empty()) {
            LevelCounters* p = g_pin_free.back();
            g_pin_free.pop_back();
            return p;
        }
    }
    LevelCounters* p = nullptr;
    if (hipHostMalloc((void**)&p, sizeof(LevelCounters) * kMaxLevelsPinned, hipHostMallocDefault) !=
        hipSuccess)
        return nullptr;
    return p;
}
static void pinned_levels_put(LevelCounters* p) {
    std::lock_guard<std::mutex> lk(g_pin_mu);
    g_pin_free.push_back(p);
}


// Most tree levels a solve may have (the reference has no bound; a deeper tree is
// reported as a device failure).
constexpr int kMaxLevels = kMaxLevelsPinned - 2;

// Grid of a grid-stride kernel for an item bound: enough workgroups for the items
// (items_per_wg each), at most cap.
static unsigned grid_for(long long items, int items_per_wg, unsigned cap) {
    const long long g = (items + items_per_wg - 1) / items_per_wg;
    return (unsigned)std::max<long long>(1, std::min<long long>(g, (long long)cap));
}
constexpr unsigned kGridCap = 8192;

// Device memory and pinned host buffers that a captured solve owns: a captured graph
// copies its descriptors from pinned memory and its results into pinned memory, and
// works in a block nothing else uses.
struct GraphWs {
    char* base = nullptr;
    size_t cap = 0;
    MatDesc* hmd = nullptr;        // descriptors (the copy source of the graph)
    int* horder = nullptr;
    MatState* hms = nullptr;       // states after a replay
    LevelCounters* hpin = nullptr; // level counters after a replay
    int* hflag = nullptr;          // the representation pool's overflow flag
    int nmat = 0;
};

// Options and results of one pass of the pipeline.
int small_release(SmallJob& job, hipStream_t after);

struct PipeOpts {
    int fixed_levels = -1;   // >= 0: exactly that many tree levels, slots from one pool
    size_t pool_cap = 0;     // doubles of that pool
    GraphWs* gw = nullptr;   // set while capturing: gw's block and buffers, no waits
    int nwaves = 0;          // > 0: singleton waves to size for (a capture repeats a run)
    // the counts of every level of an earlier solve of the same problems (a fixed walk
    // sizes its grids with them; the grid-stride kernels are right with any grid)
    std::vector<LevelCounters> hint;
    int hint_blocks = 0;
    size_t pool_need = 0;    // results: doubles of slots used, bytes of device memory,
    size_t ws_need = 0;      //   levels walked, singleton waves sized for
    int levels = 0;
    int nwaves_used = 0;
    SmallJob* small = nullptr;   // small-route work to start once the roots are refined
};

// One pass of the pipeline for the problems described by md (host copy).  Every launch
// is sized on the device: the kernels read their item counts from device memory and
// walk them with grid-stride loops, so the host never waits for the device inside the
// solve.  It enqueues the cluster work of the tree levels up to kAhead levels ahead of
// the last level whose counts it has seen (a level with no clusters ends the walk;
// levels enqueued past it find nothing to do), and the singleton phase of a level once
// it knows that the level has singletons.  When `fixed_levels` >= 0 it enqueues
// exactly that many levels and every singleton phase, without looking at any count,
// which is what a captured graph replays.  Returns 0, 2 (device failure) or 3 (the
// representation pool was too small: *pool_need holds the doubles needed).
static int pipeline_once(std::vector<MatDesc>& md, std::vector<MatState>& ms, hipStream_t st,
                         PipeOpts& o) {
    const int fixed_levels = o.fixed_levels;
    const size_t pool_cap = o.pool_cap;
    GraphWs* const gw = o.gw;
    const bool capturing = gw != nullptr;
    configure_pool_once();
    MR3_HOST_MARK("begin", -1);
    const mr3_tuning tune = tuning_snapshot();
    const int nmat = (int)md.size();
    long long ntot = 0;
    int nmax = 1;
    bool any_v = false, any_n = false, any_part = false;
    // Matrices take their slots of the eigenvalue arrays in order of decreasing size, so
    // that the lanes of a wave, which take consecutive slots, mostly belong to matrices of
    // about the same size and sweep about the same number of rows.  Results do not
    // depend on the order (each matrix is solved on its own slots).
    std::vector<int> order(nmat);
    for (int k = 0; k < nmat; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return md[a].n > md[b].n; });
    for (int k : order) {
        MatDesc& d = md[k];
        d.off = (int)ntot;
        ntot += d.n;
        nmax = std::max(nmax, d.n);
        any_v |= (d.jobz == kV);
        any_n |= (d.jobz == kN);
        any_part |= (d.range != kAll && d.jobz != kC);
    }
    if (ntot > 0x7fffffff / 2) MR3_FAIL();
    const int N = (int)ntot;
    const bool fixed = fixed_levels >= 0;
    if (capturing && (!fixed || gw->nmat != nmat)) MR3_FAIL();
    Alloc A = capturing ? Alloc(st, gw->base, gw->cap) : Alloc(st);
    CallEvents ce;
    ce.on = !capturing && g_prof.on.load() != 0;
    MatDesc* dmd = A.get<MatDesc>(nmat);
    MatState* dms = A.get<MatState>(nmat);
    int* dorder = A.get<int>(nmat);
    double* D = A.get_swept(N);
    double* E = A.get_swept(N);
    double* E2 = A.get_swept(N);
    double* gers = A.get<double>(2 * (size_t)N);
    double* Dcopy = A.get_swept(N);
    double* E2copy = A.get_swept(N);
    double* DLr = A.get_swept(N);
    double* DLLr = A.get_swept(N);
    double* W = A.get<double>(N);
    double* Werr = A.get<double>(N);
    double* Wgap = A.get<double>(N);
    double* Wsh = A.get<double>(N);
    double* Wcoarse = A.get<double>(N);
    double* Wtmp = A.get<double>(N);
    double* Werrtmp = A.get<double>(N);
    int* iblock = A.get<int>(N);
    int* Windex = A.get<int>(N);
    int* Zidx = A.get<int>(N);
    int* smat = A.get<int>(N);
    int* isplit = A.get<int>(N);
    BlockRec* blk = A.get<BlockRec>(N);
    int* tcount = A.get<int>(N);
    int* plist = A.get<int>(N);    // blocks of the shared top of the coarse bisection
    int* dbl = A.get<int>(N);      // blocks of the root representations
    int* lcount = A.get<int>(2);   // their lengths
    if (!A.ok) MR3_FAIL();
    const MatDesc* src_md = md.data();
    const int* src_order = order.data();
    if (capturing) {   // a graph copies from memory that lives as long as it does
        memcpy(gw->hmd, md.data(), sizeof(MatDesc) * nmat);
        memcpy(gw->horder, order.data(), sizeof(int) * nmat);
        src_md = gw->hmd;
        src_order = gw->horder;
    }
    MR3_CHECK(hipMemcpyAsync(dmd, src_md, sizeof(MatDesc) * nmat, hipMemcpyHostToDevice, st));
    MR3_CHECK(hipMemcpyAsync(dorder, src_order, sizeof(int) * nmat, hipMemcpyHostToDevice, st));

    LAUNCH(k_prep, dim3(nmat), dim3(256), dmd, dms, D, E, E2, gers, Dcopy,
                       E2copy, isplit, blk, W, Werr, Wcoarse, iblock, Windex, smat);
    if (any_part) LAUNCH(k_block_bracket, dim3(nmat), dim3(PREP_NT), dmd, dms, D, E2, blk);
    LAUNCH(k_block_list, dim3(1), dim3(kListNT), dmd, dms, dorder, nmat, 0, plist, lcount);
    LAUNCH(k_coarse_tree, dim3(grid_for(N, 1, 2048)), dim3(256), dmd, dms, D, E2, blk, smat, plist,
           lcount, tcount);
    LAUNCH(k_coarse, dim3(cdiv(N, 256)), dim3(256), dmd, dms, D, E2, blk, smat,
                       iblock, Windex, tcount, W, Werr, Wcoarse, N);
    LAUNCH(k_discard, dim3(cdiv(nmat, 64)), dim3(64), dmd, dms, blk, W, Werr,
                       Wcoarse, iblock, Windex, nmat);
    LAUNCH(k_block_list, dim3(1), dim3(kListNT), dmd, dms, dorder, nmat, 1, dbl, lcount + 1);
    const int* nbl_p = lcount + 1;
    LAUNCH(k_root, dim3(grid_for(N, 1, kGridCap)), dim3(64), dmd, dms, blk, smat, dbl, D, E, DLr,
           DLLr, W, Werr, Wgap, nbl_p);
    LAUNCH(k_refine_root, dim3(cdiv(N, 256)), dim3(256), dmd, dms, blk, smat,
                       iblock, Windex, D, DLLr, W, Werr, Wgap, Wtmp, Werrtmp, N);
    // the small route's smaller matrices start now: the coarse bisection and the root
    // refinement, which every later step of the pipeline waits for, have had the device
    // to themselves (apart from the small route's largest matrices)
    if (o.small && !capturing && small_release(*o.small, st)) MR3_FAIL();
    LAUNCH(k_gaps_root, dim3(cdiv(N, 256)), dim3(256), dmd, dms, blk, smat,
                       iblock, Wtmp, Werrtmp, W, Werr, Wgap, N);
    if (any_n)
        LAUNCH(k_shift_n, dim3(cdiv(N, 256)), dim3(256), dmd, dms, blk, smat,
                           iblock, W, N);
    MR3_CHECK(hipGetLastError());
    int levels_done = 0;
    size_t pool_used = 0;   // slots of the levels (eager walk)
    int* pool_flag = nullptr;
    unsigned long long* bump = nullptr;
    // pinned copies of each level's counters (the host reads them without waiting)
    LevelCounters* hpin = nullptr;
    if (any_v) {
        LAUNCH(k_zindex, dim3(nmat), dim3(256), dmd, dms, blk, W, iblock, Windex,
                           Zidx);
        // task lists: singletons (<= N), clusters (<= N/2 + 1), eigen entries (<= N)
        const int capS = N, capC = N / 2 + 1, capE = N;
        // wave tasks of a level: ceil(count / 64) per block, at most N blocks
        const int capT = N / 64 + N + 1;
        // singleton lists of four consecutive levels: the side streams may still work on
        // levels L - 2 .. L while the main stream classifies level L + 1
        constexpr int kLB = 4;
        STask* Sl2[kLB];
        int* bcnt2[kLB];
        for (int k = 0; k < kLB; ++k) {
            Sl2[k] = A.get<STask>(capS);
            bcnt2[k] = A.get<int>(N);
        }
        ClTask* Cl[2] = {A.get<ClTask>(capC), A.get<ClTask>(capC)};
        int* es[2] = {A.get<int>(capE), A.get<int>(capE)};
        int* ec[2] = {A.get<int>(capE), A.get<int>(capE)};
        // counters of the level being worked on and of the next one (each level's
        // classification writes the next level's)
        LevelCounters* lc = A.get<LevelCounters>(2);
        bump = A.get<unsigned long long>(1);
        pool_flag = A.get<int>(1);
        // fixed levels: one pool for the slots of all levels; otherwise each level gets
        // slots of the size its counts ask for
        double* pool = fixed ? A.get_swept(pool_cap) : nullptr;
        double* spare = A.get_swept(4 * (size_t)nmax);
        // singleton iterations, one set per side stream: the first iteration's tasks,
        // per-singleton state, two generations of compacted lists (S entries per
        // block) with their tasks and counters, the bisection list
        struct SSet {
            WaveTask* wtasks;
            LevelCounters* scnt;
            SState* ST;
            WaveTask* tit[2];
            int* sit[2];
            LevelCounters* cit[2];
            int* sbc;
            int* BI;
            int* DI;          // dense list of the tail kernel
            int* DIb;         // the same for blocks above kTailN rows
            double* scratch;
            double* tscratch; // S and P of the large-block tail kernel, per workgroup
        } ss[kSide] = {};
        // a fixed walk (what a graph captures) takes at most two side streams: HIP runs
        // a graph's independent branches on at most four streams at once, and the graph
        // has the main stream and the small route's besides
        const int nside = fixed ? std::min(2, side_count(tune.side_streams)) : side_count(tune.side_streams);
        for (int k = 0; k < nside; ++k) {
            ss[k].wtasks = A.get<WaveTask>(capT);
            ss[k].scnt = A.get<LevelCounters>(1);
            ss[k].ST = A.get<SState>(capS);
            for (int b = 0; b < 2; ++b) {
                ss[k].tit[b] = A.get<WaveTask>(capT);
                ss[k].sit[b] = A.get<int>(N);
                ss[k].cit[b] = A.get<LevelCounters>(1);
            }
            ss[k].sbc = A.get<int>(N);
            ss[k].BI = A.get<int>(N);
            ss[k].DI = A.get<int>(N);
            ss[k].DIb = A.get<int>(N);
        }
        if (!A.ok) MR3_FAIL();
        // eigenvector scratch: two checkpoint arrays of nmax/kChunk + 2 carries per
        // lane, 64 lanes per wave (nmax bounds every block)
        const int maxbl = nmax;
        const size_t ck_per_lane = (size_t)2 * (maxbl / kChunk + 2);
        const size_t per_wave = ck_per_lane * 64 * sizeof(double);
        const size_t budget = singleton_scratch_budget() / nside;
        const size_t roomy = std::max<size_t>(singleton_waves_resident(), ((size_t)2 << 30) / per_wave);
        int max_waves = (int)std::max<size_t>(
            1, std::min(std::min<size_t>(budget / per_wave, (size_t)tune.singleton_waves), roomy));
        if (o.nwaves > 0) max_waves = o.nwaves;
        const int nwaves = std::min(max_waves, capT);
        o.nwaves_used = nwaves;
        for (int k = 0; k < nside; ++k) ss[k].scratch = A.get<double>((size_t)nwaves * ck_per_lane * 64);
        // the large-block tail kernel: S and P of up to kTailBigWG singletons at a time
        const int tstride = 2 * (maxbl + 2);
        const int tail_rows = (maxbl > kTailN) ? maxbl : kTailN;   // blocks the tail kernels take
        for (int k = 0; k < nside; ++k)
            ss[k].tscratch = (maxbl > kTailN) ? A.get<double>((size_t)kTailBigWG * tstride) : nullptr;
        hpin = capturing ? gw->hpin : pinned_levels_get();
        if (!A.ok || !hpin) {
            if (hpin && !capturing) pinned_levels_put(hpin);
            MR3_FAIL();
        }
        struct PinFree {
            LevelCounters* p;
            ~PinFree() { if (p) pinned_levels_put(p); }
        } pin_free{capturing ? nullptr : hpin};
        // cleared with kernels, not memsets: a captured graph replays them
        for (int k = 0; k < kLB; ++k)
            LAUNCH(k_fill0, dim3(grid_for(N, 256, 1024)), dim3(256), bcnt2[k], (long long)N);
        for (int k = 0; k < nside; ++k)
            LAUNCH(k_fill0, dim3(grid_for(N, 256, 1024)), dim3(256), ss[k].sbc, (long long)N);
        LAUNCH(k_zero, dim3(1), dim3(64), (int*)bump, 2, (int*)lc,
               (int)(2 * sizeof(LevelCounters) / sizeof(int)));
        LAUNCH(k_zero, dim3(1), dim3(64), pool_flag, 1, (int*)nullptr, 0);
        TreeOut out;
        out.S = Sl2[0]; out.C = Cl[0]; out.eig_slot = es[0]; out.eig_cl = ec[0]; out.cnt = lc;
        out.bcnt = bcnt2[0];
        out.halt = pool_flag;
        out.capS = capS; out.capC = capC; out.capEig = capE;
        LAUNCH(k_classify_root, dim3(grid_for(N, 1, kGridCap)), dim3(64), dmd, dms, blk,
                           smat, dbl, gers, D, E, DLr, DLLr, W, Werr, Wgap, Wsh, out, nbl_p);
        MR3_CHECK(hipGetLastError());
        hipStream_t s2p[kSide] = {};
        for (int k = 0; k < nside; ++k) {
            s2p[k] = side_stream(k);
            if (!s2p[k]) MR3_FAIL();
        }
        // events: end of the singleton phase that last read list buffer k; the lists of
        // level L ready on the main stream; level L's counts copied to pinned memory
        struct EvSet {
            std::vector<hipEvent_t> e;
            ~EvSet() { for (hipEvent_t x : e) event_pool().put(x); }
            hipEvent_t get() { e.push_back(event_pool().get()); return e.back(); }
        } evs;
        hipEvent_t evS[kLB] = {};
        bool have_evS[kLB] = {};
        for (int k = 0; k < kLB; ++k) evS[k] = evs.get();
        hipEvent_t evReady[kMaxLevels + 1] = {}, evCnt[kMaxLevels + 1] = {};
        struct SideSync {   // no buffer or event goes away while a side stream may use it
            const hipStream_t* s;
            bool capturing;
            ~SideSync() {
                if (capturing) return;   // (the capture joined them to the main stream)
                for (int k = 0; k < kSide; ++k)
                    if (s[k]) (void)hipStreamSynchronize(s[k]);
            }
        } side_sync{s2p, capturing};
        // level L's counts, copied to pinned memory once its lists are complete
        auto publish = [&](int L, const LevelCounters* src) -> int {
            evReady[L] = evs.get();
            MR3_CHECK(hipEventRecord(evReady[L], st));
            MR3_CHECK(hipMemcpyAsync(&hpin[L], src, sizeof(LevelCounters), hipMemcpyDeviceToHost, st));
            evCnt[L] = evs.get();
            MR3_CHECK(hipEventRecord(evCnt[L], st));
            return 0;
        };
        // the number of blocks, ahead of level 0's counts (it sizes the planning launches)
        LevelCounters& hblocks = hpin[kMaxLevelsPinned - 1];
        MR3_CHECK(hipMemcpyAsync(&hblocks.nS, nbl_p, sizeof(int), hipMemcpyDeviceToHost, st));
        if (publish(0, lc)) MR3_FAIL();
        // what the host knows: the counts of levels 0 .. known - 1
        int known = 0;
        auto learn = [&](int upto, bool wait) -> int {   // reads the counts that arrived
            while (known <= upto) {
                if (wait) MR3_CHECK(hipEventSynchronize(evCnt[known]));
                else if (hipEventQuery(evCnt[known]) != hipSuccess) break;
                if (hpin[known].err) MR3_FAIL();
                ++known;
            }
            return 0;
        };
        // the counts of level L when the host has them: this solve's, or in a fixed walk
        // those of the earlier solve the caller passed
        auto counts = [&](int L) -> const LevelCounters* {
            if (!fixed) return known > L ? &hpin[L] : nullptr;
            return L < (int)o.hint.size() ? &o.hint[L] : nullptr;
        };
        // item bound of level L: the eigen entries of the latest level seen at or before it
        auto bound = [&](int L) -> long long {
            if (fixed) {
                const LevelCounters* c = counts(L);
                return c ? std::max(1, c->nEig + c->nS) : N;
            }
            for (int j = std::min(L, known - 1); j >= 0; --j) return std::max(1, hpin[j].nEig + hpin[j].nS);
            return N;
        };
        const long long gbud = tune.group_lanes;
        const int rrf_max = tune.rrf_group_max;
        // the cluster work of level L and the classification into level L + 1 (main stream)
        auto main_level = [&](int L, double* arena, size_t arena_cap) -> int {
            const int cur = L & 1, nxt = cur ^ 1;
            const long long b = bound(L);
            const long long bc = std::min<long long>(capC, b / 2 + 1);
            LevelCounters* lcur = lc + cur;
            LevelCounters* lnxt = lc + nxt;
            LAUNCH(k_zero, dim3(1), dim3(64), (int*)lnxt,
                   (int)(sizeof(LevelCounters) / sizeof(int)), (int*)nullptr, 0);
            if (!fixed)   // a slot of the level's own: the bump counter starts again
                LAUNCH(k_zero, dim3(1), dim3(64), (int*)bump, 2, (int*)nullptr, 0);
            LAUNCH(k_arena_assign, dim3(grid_for(bc, 64, kGridCap)), dim3(64), Cl[cur], &lcur->nC, arena,
                   bump, (unsigned long long)arena_cap, spare, pool_flag);
            LAUNCH(k_endref, dim3(grid_for(2 * bc, 64, kGridCap)), dim3(64), Cl[cur], &lcur->nC, Windex,
                   Wsh, Werr, dms, Wtmp, Werrtmp, gbud);
            LAUNCH(k_endref_grp, dim3(grid_for(std::min<long long>(gbud, 64 * 2 * bc), 64, kGridCap)),
                   dim3(64), Cl[cur], &lcur->nC, gbud, Windex, Wsh, Werr, dms, Wtmp, Werrtmp);
            // four lanes per cluster unless the clusters alone fill the device
            LAUNCH(k_dlarrf<1>, dim3(grid_for(bc, 64, kGridCap)), dim3(64), Cl[cur], &lcur->nC,
                   rrf_max, Wsh, Werr, Wgap, Wtmp, Werrtmp, dms);
            LAUNCH(k_dlarrf<4>, dim3(grid_for(std::min<long long>(bc, rrf_max), 16, kGridCap)), dim3(64),
                   Cl[cur], &lcur->nC, rrf_max, Wsh, Werr, Wgap, Wtmp, Werrtmp, dms);
            LAUNCH(k_rrf_store, dim3(grid_for(std::min<long long>(bc, rrf_max), 64, kGridCap)), dim3(64),
                   Cl[cur], &lcur->nC, rrf_max, dms);
            LAUNCH(k_newrrr_fin, dim3(grid_for(bc, 1, kGridCap)), dim3(64), Cl[cur], &lcur->nC, Wsh,
                   Werr);
            LAUNCH(k_refine_cl, dim3(grid_for(b, 64, kGridCap)), dim3(64), Cl[cur], es[cur], ec[cur],
                   &lcur->nEig, Windex, Wsh, Werr, Wgap, dms, Wtmp, Werrtmp, gbud);
            LAUNCH(k_refine_cl_grp, dim3(grid_for(std::min<long long>(gbud, 64 * b), 64, kGridCap)),
                   dim3(64), Cl[cur], es[cur], ec[cur], &lcur->nEig, gbud, Windex, Wsh, Werr, Wgap,
                   dms, Wtmp, Werrtmp);
            LAUNCH(k_gaps_cl, dim3(grid_for(b, 64, kGridCap)), dim3(64), Cl[cur], es[cur], ec[cur],
                   &lcur->nEig, Wtmp, Werrtmp, Wsh, Werr, Wgap, W);
            // the next level's singleton list reuses the buffers of the phase four
            // levels back, which a side stream may still be reading
            const int lbn = (L + 1) % kLB;
            if (have_evS[lbn]) MR3_CHECK(hipStreamWaitEvent(st, evS[lbn], 0));
            TreeOut o = out;
            o.C = Cl[nxt]; o.eig_slot = es[nxt]; o.eig_cl = ec[nxt];
            o.S = Sl2[lbn];
            o.bcnt = bcnt2[lbn];
            o.cnt = lnxt;
            LAUNCH(k_classify, dim3(grid_for(bc, 64, kGridCap)), dim3(64), Cl[cur], &lcur->nC, Wgap,
                   Wsh, o);
            MR3_CHECK(hipGetLastError());
            return publish(L + 1, lnxt);
        };
        // the singletons of level L on a side stream, concurrently with the main stream
        // and with the other side streams' earlier levels; every launch sizes itself on
        // the device
        auto side_level = [&](int L) -> int {
            const int p = L % nside;
            hipStream_t s2 = s2p[p];
            SSet& q = ss[p];
            const int lb = L % kLB;
            STask* Sl = Sl2[lb];
            int* bcnt = bcnt2[lb];
            // the level's lists, representations and eigenvalues are complete on the main
            // stream when its counts are published; the main stream's later work touches
            // other slots
            MR3_CHECK(hipStreamWaitEvent(s2, evReady[L], 0));
            // grids: from the level's counts when the host has them, else from bounds
            const LevelCounters* c = counts(L);
            const long long nS = c ? c->nS : N;
            const long long nb = !fixed ? (known > 0 ? hblocks.nS : N) : (o.hint_blocks > 0 ? o.hint_blocks : N);
            const int g1 = (int)std::min<long long>(capT, nS / 64 + nb + 1);   // wave tasks
            const unsigned gT = grid_for(g1, 1, kGridCap);
            const unsigned gw = (unsigned)std::min(nwaves, g1);
            const unsigned gB = grid_for(nb, 64, 1024);
            const unsigned gTail = (unsigned)std::min<long long>(nS, 1024);
            const unsigned gTailB = (unsigned)std::min<long long>(nS, kTailBigWG);
            // first iteration over the level's singleton lists (64 per task,
            // longest blocks first)
            LAUNCH_ON(s2, k_zero, dim3(1), dim3(64), (int*)q.scnt,
                      (int)(sizeof(LevelCounters) / sizeof(int)), (int*)nullptr, 0);
            LAUNCH_ON(s2, k_plan_count, dim3(gB), dim3(64), dbl, nbl_p, blk, bcnt, 64,
                      q.scnt, q.cit[1]);
            LAUNCH_ON(s2, k_plan_s, dim3(gB), dim3(64), dbl, nbl_p, smat, dmd, blk, bcnt,
                      64, q.wtasks, capT, q.scnt);
            LAUNCH_ON(s2, k_s_iter, dim3(gw), dim3(64), q.wtasks, q.scnt, 1, (const int*)nullptr,
                      Sl, dmd, dms, Windex, Zidx, W, Wsh, Werr, Wgap, q.ST, q.scratch, maxbl,
                      tune.tail_singletons, tail_rows);
            // further iterations on the compacted singletons that go on (the
            // reference's loop runs at most kMaxIterRQI dlar1v calls plus the
            // best-residual step)
            const WaveTask* ptasks = q.wtasks;
            const LevelCounters* pcnt = q.scnt;
            const int* psi = nullptr;
            int pfirst = 1;
            for (int it = 1; it <= kMaxIterRQI + 1; ++it) {
                const int b = it & 1;
                // q.cit[b] was cleared by the previous iteration's k_plan_count
                LAUNCH_ON(s2, k_s_collect, dim3(gT), dim3(64), ptasks, pcnt, pfirst, psi, Sl,
                          q.ST, q.sit[b], q.sbc, q.BI, q.DI, q.DIb, q.cit[b]);
                LAUNCH_ON(s2, k_plan_count, dim3(gB), dim3(64), dbl, nbl_p, blk, q.sbc,
                          64, q.cit[b], it < kMaxIterRQI + 1 ? q.cit[b ^ 1] : (LevelCounters*)nullptr);
                LAUNCH_ON(s2, k_plan_s, dim3(gB), dim3(64), dbl, nbl_p, smat, dmd, blk,
                          q.sbc, 64, q.tit[b], capT, q.cit[b]);
                LAUNCH_ON(s2, k_s_bisect, dim3(std::min(gw, 2048u)), dim3(64), q.BI, q.cit[b], Sl,
                          dms, Windex, q.ST);
                LAUNCH_ON(s2, k_s_iter, dim3(gw), dim3(64), q.tit[b], q.cit[b], 0, q.sit[b], Sl,
                          dmd, dms, Windex, Zidx, W, Wsh, Werr, Wgap, q.ST, q.scratch, maxbl,
                          tune.tail_singletons, tail_rows);
                LAUNCH_ON(s2, k_s_tail<false>, dim3(gTail), dim3(128), q.DI, q.cit[b],
                          Sl, dmd, dms, Windex, Zidx, W, Wsh, Werr, Wgap, q.ST,
                          kMaxIterRQI + 2 - it, tune.tail_singletons, (double*)nullptr, 0);
                if (maxbl > kTailN)
                    LAUNCH_ON(s2, k_s_tail<true>, dim3(gTailB), dim3(128), q.DIb,
                              q.cit[b], Sl, dmd, dms, Windex, Zidx, W, Wsh, Werr, Wgap, q.ST,
                              kMaxIterRQI + 2 - it, tune.tail_singletons, q.tscratch, tstride);
                ptasks = q.tit[b];
                pcnt = q.cit[b];
                psi = q.sit[b];
                pfirst = 0;
            }
            LAUNCH_ON(s2, k_s_normalize, dim3(gT), dim3(64), q.wtasks, q.scnt, Sl, dmd, dms, Zidx, q.ST);
            MR3_CHECK(hipGetLastError());
            MR3_CHECK(hipEventRecord(evS[lb], s2));
            have_evS[lb] = true;
            return 0;
        };
        int mainL = 0;
        MR3_HOST_MARK("walk", -1);
        if (fixed) {
            // the clusters of levels 0 .. fixed_levels - 1 and the singletons of levels
            // 0 .. fixed_levels (the last level has singletons only)
            const int nl = std::min(fixed_levels, kMaxLevels);
            for (int L = 0; L <= nl; ++L) {
                if (L < nl && main_level(L, pool, pool_cap)) MR3_FAIL();
                if (side_level(L)) MR3_FAIL();
            }
            mainL = nl;
        } else {
            // level L's counts arrive (an event, not a stream wait: the side streams and
            // the copy engines go on) before its work is enqueued: they size the slots
            // of the level's child representations
            for (int L = 0;; ++L) {
                if (learn(L, true)) MR3_FAIL();
                MR3_HOST_MARK("learned", L);
                if (hpin[L].nS > 0 && side_level(L)) MR3_FAIL();
                MR3_HOST_MARK("side", L);
                if (hpin[L].nC == 0) break;
                if (L >= kMaxLevels) MR3_FAIL();   // deeper than any tree the solver allows
                const size_t need = (size_t)hpin[L].arena;
                double* arena = A.get_swept(need);
                if (!A.ok) MR3_FAIL();
                pool_used += need;
                if (main_level(L, arena, need)) MR3_FAIL();
                MR3_HOST_MARK("main", L);
                mainL = L + 1;
            }
        }
        for (int k = 0; k < kLB; ++k)
            if (have_evS[k]) MR3_CHECK(hipStreamWaitEvent(st, evS[k], 0));
        levels_done = mainL;
    }
    LAUNCH(k_rrj, dim3(cdiv(N, 256)), dim3(256), dmd, dms, blk, smat, iblock,
                       Windex, Dcopy, E2copy, W, Werr, N);
    // (tcount and isplit are free by now: the reordering's permutation and marks)
    LAUNCH(k_finalize, dim3(nmat), dim3(256), dmd, dms, W, Zidx, blk, iblock,
                       Windex, tcount, isplit, Wcoarse);
    MR3_CHECK(hipGetLastError());
    o.levels = levels_done;
    o.ws_need = A.need;
    if (capturing) {   // results into the graph's pinned buffers; the caller waits
        MR3_CHECK(hipMemcpyAsync(gw->hms, dms, sizeof(MatState) * nmat, hipMemcpyDeviceToHost, st));
        if (pool_flag)
            MR3_CHECK(hipMemcpyAsync(gw->hflag, pool_flag, sizeof(int), hipMemcpyDeviceToHost, st));
        return 0;
    }
    ms.resize(nmat);
    MR3_CHECK(hipMemcpyAsync(ms.data(), dms, sizeof(MatState) * nmat, hipMemcpyDeviceToHost, st));
    int hflag = 0;
    unsigned long long hbump = 0;
    if (pool_flag) {
        MR3_CHECK(hipMemcpyAsync(&hflag, pool_flag, sizeof(int), hipMemcpyDeviceToHost, st));
        MR3_CHECK(hipMemcpyAsync(&hbump, bump, sizeof(hbump), hipMemcpyDeviceToHost, st));
    }
    MR3_CHECK(hipStreamSynchronize(st));
    ce.collect();
    o.pool_need = fixed ? (size_t)hbump : pool_used;
    if (hpin && !fixed) {
        o.hint.assign(hpin, hpin + std::min(levels_done, kMaxLevels) + 1);
        o.hint_blocks = hpin[kMaxLevelsPinned - 1].nS;
    }
    if (hflag) return 3;
    if (fixed && hpin && levels_done <= kMaxLevels && hpin[levels_done].nC > 0)
        return 4;   // the tree is deeper than the levels walked
    if (hpin)
        for (int j = 0; j <= levels_done && j <= kMaxLevels; ++j)
            if (hpin[j].err) MR3_FAIL();
    return 0;
}

// Runs the whole pipeline for the problems described by md (host copy); the device
// outputs land in the caller's arrays named in md; states are copied back to ms.
int run_pipeline(std::vector<MatDesc>& md, std::vector<MatState>& ms, hipStream_t st,
                 SmallJob* small) {
    PipeOpts o;
    o.small = small;
    const int rc = pipeline_once(md, ms, st, o);
    return rc == 0 ? 0 : 2;   // (the eager walk sizes every level's slots: no overflow)
}

// ---------------------------------------------------------------------------------
// The small-matrix route (n < 500, the reference's dstemr): k_small on streams of its
// own, started before the large problems' pipeline and collected after it, so that both
// run on the device at the same time.  Its largest matrices (order >= small_first)
// start at once: each is one long sequential chain on one lane that would otherwise end
// last.  The others, many short solves that fill the device, start when the pipeline
// has refined its root eigenvalues (small_release): before that the coarse bisection
// and the root refinement, on which every later pipeline step waits, have the device.
// ---------------------------------------------------------------------------------
static hipStream_t small_stream(int which) {
    thread_local hipStream_t s[2] = {nullptr, nullptr};
    thread_local bool tried[2] = {false, false};
    if (!tried[which]) {
        tried[which] = true;
        if (hipStreamCreateWithFlags(&s[which], hipStreamNonBlocking) != hipSuccess) s[which] = nullptr;
    }
    return s[which];
}

struct SmallJob::Impl {
    hipStream_t s = nullptr, s2 = nullptr;
    char* buf = nullptr;
    MatState* dms = nullptr;
    hipEvent_t ready = nullptr;     // the descriptors are on the device
    hipEvent_t done2 = nullptr;     // the deferred launch has ended
    int nmat = 0;
    int nfirst = 0;                 // matrices launched at once (the largest)
    bool pending = false;           // the others not launched yet
    MatDesc* dmd = nullptr;
    int* dlist = nullptr;
    long long* doff = nullptr;
    char* slices = nullptr;
    std::vector<int> list;
    std::vector<long long> off;
    std::vector<int> nlist;         // order of the matrix of each list entry
    CallEvents ce;
    ~Impl() {   // a job that failed before small_end
        if (buf) {
            (void)hipStreamSynchronize(s);
            if (s2) (void)hipStreamSynchronize(s2);
            (void)hipFree(buf);
        }
        if (ready) event_pool().put(ready);
        if (done2) event_pool().put(done2);
    }
};

SmallJob::SmallJob() : p(new Impl) {}
SmallJob::~SmallJob() { delete p; }

// k_small on list entries [first, first + count) on stream s
static int small_launch(SmallJob::Impl& J, hipStream_t s, int first, int count) {
    if (count <= 0) return 0;
    CallEvents& ce = J.ce;
    const size_t lds = sizeof(double) * (size_t)mr3sm::lds_doubles(std::max(1, J.nlist[first]));
    hipEvent_t e0_ = nullptr, e1_ = nullptr;
    if (ce.on) {
        e0_ = event_pool().get();
        e1_ = event_pool().get();
        (void)hipEventRecord(e0_, s);
    }
    hipLaunchKernelGGL(k_small, dim3(count), dim3(kSmallNT), lds, s, J.dmd, J.dms, J.dlist + first,
                       J.doff + first, J.slices);
    if (ce.on) {
        (void)hipEventRecord(e1_, s);
        ce.ev.push_back({"k_small", {e0_, e1_}});
    }
    MR3_CHECK(hipGetLastError());
    return 0;
}

// defer: a pipeline runs beside the route and calls small_release when its roots are
// refined; otherwise every matrix starts at once
int small_begin(SmallJob& job, const std::vector<MatDesc>& md, hipStream_t st, bool defer) {
    SmallJob::Impl& J = *job.p;
    const int nm = (int)md.size();
    J.nmat = nm;
    if (nm == 0) return 0;
    J.s = small_stream(0);
    if (!J.s) return 2;
    configure_pool_once();
    // largest first: the longest solves start first
    J.list.resize(nm);
    for (int i = 0; i < nm; ++i) J.list[i] = i;
    std::stable_sort(J.list.begin(), J.list.end(), [&](int a, int b) { return md[a].n > md[b].n; });
    J.off.resize(nm);
    J.nlist.resize(nm);
    long long tot = 0;
    for (int j = 0; j < nm; ++j) {
        J.off[j] = tot;
        J.nlist[j] = md[J.list[j]].n;
        tot += small_slice_bytes(md[J.list[j]].n);
    }
    const int first_n = tuning_snapshot().small_first;
    J.nfirst = nm;
    if (defer && first_n > 0) {
        J.nfirst = 0;
        while (J.nfirst < nm && J.nlist[J.nfirst] >= first_n) ++J.nfirst;
    }
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t bmd = al(sizeof(MatDesc) * nm), bms = al(sizeof(MatState) * nm);
    const size_t bl = al(sizeof(int) * nm), bo = al(sizeof(long long) * nm);
    hipEvent_t e0 = event_pool().get();
    MR3_CHECK(hipEventRecord(e0, st));
    MR3_CHECK(hipStreamWaitEvent(J.s, e0, 0));
    event_pool().put(e0);
    MR3_CHECK(hipMallocAsync((void**)&J.buf, bmd + bms + bl + bo + (size_t)tot, J.s));
    J.dmd = (MatDesc*)J.buf;
    J.dms = (MatState*)(J.buf + bmd);
    J.dlist = (int*)(J.buf + bmd + bms);
    J.doff = (long long*)(J.buf + bmd + bms + bl);
    J.slices = J.buf + bmd + bms + bl + bo;
    MR3_CHECK(hipMemcpyAsync(J.dmd, md.data(), sizeof(MatDesc) * nm, hipMemcpyHostToDevice, J.s));
    MR3_CHECK(hipMemcpyAsync(J.dlist, J.list.data(), sizeof(int) * nm, hipMemcpyHostToDevice, J.s));
    MR3_CHECK(hipMemcpyAsync(J.doff, J.off.data(), sizeof(long long) * nm, hipMemcpyHostToDevice, J.s));
    J.ready = event_pool().get();
    MR3_CHECK(hipEventRecord(J.ready, J.s));
    J.ce.on = g_prof.on.load() != 0;
    if (small_launch(J, J.s, 0, J.nfirst)) return 2;
    J.pending = J.nfirst < nm;
    return 0;
}

// Starts the deferred matrices once the work enqueued on `after` so far is done.
int small_release(SmallJob& job, hipStream_t after) {
    SmallJob::Impl& J = *job.p;
    if (!J.pending) return 0;
    J.pending = false;
    J.s2 = small_stream(1);
    if (!J.s2) return 2;
    hipEvent_t ea = event_pool().get();
    MR3_CHECK(hipEventRecord(ea, after));
    MR3_CHECK(hipStreamWaitEvent(J.s2, J.ready, 0));
    MR3_CHECK(hipStreamWaitEvent(J.s2, ea, 0));
    event_pool().put(ea);
    if (small_launch(J, J.s2, J.nfirst, J.nmat - J.nfirst)) return 2;
    J.done2 = event_pool().get();
    MR3_CHECK(hipEventRecord(J.done2, J.s2));
    return 0;
}

int small_end(SmallJob& job, std::vector<MatState>& ms, hipStream_t st) {
    SmallJob::Impl& J = *job.p;
    if (J.nmat == 0) return 0;
    if (J.pending && small_release(job, J.s)) return 2;   // no pipeline released them
    if (J.done2) MR3_CHECK(hipStreamWaitEvent(J.s, J.done2, 0));
    ms.resize(J.nmat);
    MR3_CHECK(hipMemcpyAsync(ms.data(), J.dms, sizeof(MatState) * J.nmat, hipMemcpyDeviceToHost, J.s));
    MR3_CHECK(hipStreamSynchronize(J.s));
    J.ce.collect();
    (void)hipFreeAsync(J.buf, J.s);
    J.buf = nullptr;
    MR3_CHECK(hipEventRecord(J.ready, J.s));
    MR3_CHECK(hipStreamWaitEvent(st, J.ready, 0));
    event_pool().put(J.ready);
    J.ready = nullptr;
    if (J.done2) {
        event_pool().put(J.done2);
        J.done2 = nullptr;
    }
    return 0;
}

// ---------------------------------------------------------------------------------
// Captured solves (mr3_graph_*): one hipGraph replays the whole batch, the small route
// and the pipeline with a fixed number of tree levels, in memory the graph owns.
// ---------------------------------------------------------------------------------
struct SmallGraph {   // the small route's part of a captured batch
    int nmat = 0, nmax = 1;
    char* buf = nullptr;            // descriptors, states, list, offsets, slices (device)
    MatDesc* hmd = nullptr;         // pinned copy sources and result destination
    int* hlist = nullptr;
    long long* hoff = nullptr;
    MatState* hms = nullptr;
    size_t bmd = 0, bms = 0, bl = 0, bo = 0;
};

static void small_graph_free(SmallGraph& g) {
    if (g.buf) (void)hipFree(g.buf);
    if (g.hmd) (void)hipHostFree(g.hmd);
    if (g.hlist) (void)hipHostFree(g.hlist);
    if (g.hoff) (void)hipHostFree(g.hoff);
    if (g.hms) (void)hipHostFree(g.hms);
    g = SmallGraph();
}

static int small_graph_alloc(SmallGraph& g, const std::vector<MatDesc>& md) {
    const int nm = (int)md.size();
    g.nmat = nm;
    if (nm == 0) return 0;
    std::vector<int> list(nm);
    for (int i = 0; i < nm; ++i) list[i] = i;
    std::stable_sort(list.begin(), list.end(), [&](int a, int b) { return md[a].n > md[b].n; });
    auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    g.bmd = al(sizeof(MatDesc) * nm); g.bms = al(sizeof(MatState) * nm);
    g.bl = al(sizeof(int) * nm); g.bo = al(sizeof(long long) * nm);
    MR3_CHECK(hipHostMalloc((void**)&g.hmd, sizeof(MatDesc) * nm, hipHostMallocDefault));
    MR3_CHECK(hipHostMalloc((void**)&g.hlist, sizeof(int) * nm, hipHostMallocDefault));
    MR3_CHECK(hipHostMalloc((void**)&g.hoff, sizeof(long long) * nm, hipHostMallocDefault));
    MR3_CHECK(hipHostMalloc((void**)&g.hms, sizeof(MatState) * nm, hipHostMallocDefault));
    long long tot = 0;
    for (int j = 0; j < nm; ++j) {
        g.hlist[j] = list[j];
        g.hoff[j] = tot;
        tot += small_slice_bytes(md[list[j]].n);
        g.nmax = std::max(g.nmax, md[list[j]].n);
    }
    memcpy(g.hmd, md.data(), sizeof(MatDesc) * nm);
    MR3_CHECK(hipMalloc((void**)&g.buf, g.bmd + g.bms + g.bl + g.bo + (size_t)tot));
    return 0;
}

// Enqueues the small route of a captured batch: a fork from st to the small-route
// stream, copies from and to pinned memory, one launch; *join is recorded at its end
// (st waits for it once the pipeline is enqueued, so that both run at the same time).
static int small_graph_enqueue(SmallGraph& g, hipStream_t st, hipEvent_t* join_out) {
    *join_out = nullptr;
    if (g.nmat == 0) return 0;
    hipStream_t s = small_stream(0);
    if (!s) return 2;
    const int nm = g.nmat;
    MatDesc* dmd = (MatDesc*)g.buf;
    MatState* dms = (MatState*)(g.buf + g.bmd);
    int* dlist = (int*)(g.buf + g.bmd + g.bms);
    long long* doff = (long long*)(g.buf + g.bmd + g.bms + g.bl);
    char* slices = g.buf + g.bmd + g.bms + g.bl + g.bo;
    hipEvent_t fork = event_pool().get(), join = event_pool().get();
    MR3_CHECK(hipEventRecord(fork, st));
    MR3_CHECK(hipStreamWaitEvent(s, fork, 0));
    MR3_CHECK(hipMemcpyAsync(dmd, g.hmd, sizeof(MatDesc) * nm, hipMemcpyHostToDevice, s));
    MR3_CHECK(hipMemcpyAsync(dlist, g.hlist, sizeof(int) * nm, hipMemcpyHostToDevice, s));
    MR3_CHECK(hipMemcpyAsync(doff, g.hoff, sizeof(long long) * nm, hipMemcpyHostToDevice, s));
    const size_t lds = sizeof(double) * (size_t)mr3sm::lds_doubles(g.nmax);
    hipLaunchKernelGGL(k_small, dim3(nm), dim3(kSmallNT), lds, s, dmd, dms, dlist, doff, slices);
    MR3_CHECK(hipGetLastError());
    MR3_CHECK(hipMemcpyAsync(g.hms, dms, sizeof(MatState) * nm, hipMemcpyDeviceToHost, s));
    MR3_CHECK(hipEventRecord(join, s));
    event_pool().put(fork);
    *join_out = join;
    return 0;
}

struct BatchGraph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int nprob = 0;
    std::vector<int> idxS, idxL;    // problems of the small route and of the pipeline
    SmallGraph sg;
    GraphWs gw;
    int levels = 0;
    ~BatchGraph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        small_graph_free(sg);
        if (gw.base) (void)hipFree(gw.base);
        if (gw.hmd) (void)hipHostFree(gw.hmd);
        if (gw.horder) (void)hipHostFree(gw.horder);
        if (gw.hms) (void)hipHostFree(gw.hms);
        if (gw.hpin) (void)hipHostFree(gw.hpin);
        if (gw.hflag) (void)hipHostFree(gw.hflag);
    }
};

// Captures the batch (mdS: the small route's problems, mdL: the pipeline's) as one
// graph.  The pipeline's problems are solved once as usual to learn the depth of their
// tree and the slots it needs, then once more with exactly that many levels and one
// pool (this also sizes the device memory), then captured.
int graph_create(std::vector<MatDesc>& mdS, std::vector<MatDesc>& mdL, hipStream_t st,
                 BatchGraph* g) {
    configure_pool_once();
    if (small_graph_alloc(g->sg, mdS)) return 2;
    PipeOpts ofix;
    if (!mdL.empty()) {
        std::vector<MatState> ms;
        PipeOpts oe;
        if (pipeline_once(mdL, ms, st, oe) != 0) return 2;
        ofix.fixed_levels = oe.levels;
        ofix.pool_cap = oe.pool_need + 1024;
        ofix.hint = oe.hint;
        ofix.hint_blocks = oe.hint_blocks;
        if (pipeline_once(mdL, ms, st, ofix) != 0) return 2;
        g->levels = oe.levels;
        GraphWs& w = g->gw;
        const int nm = (int)mdL.size();
        w.nmat = nm;
        w.cap = ofix.ws_need + ((size_t)1 << 20);
        MR3_CHECK(hipMalloc((void**)&w.base, w.cap));
        MR3_CHECK(hipHostMalloc((void**)&w.hmd, sizeof(MatDesc) * nm, hipHostMallocDefault));
        MR3_CHECK(hipHostMalloc((void**)&w.horder, sizeof(int) * nm, hipHostMallocDefault));
        MR3_CHECK(hipHostMalloc((void**)&w.hms, sizeof(MatState) * nm, hipHostMallocDefault));
        MR3_CHECK(hipHostMalloc((void**)&w.hpin, sizeof(LevelCounters) * kMaxLevelsPinned,
                                hipHostMallocDefault));
        MR3_CHECK(hipHostMalloc((void**)&w.hflag, sizeof(int), hipHostMallocDefault));
        *w.hflag = 0;
    }
    MR3_CHECK(hipStreamSynchronize(st));
    MR3_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed));
    hipEvent_t join = nullptr;
    int rc = small_graph_enqueue(g->sg, st, &join);
    if (rc == 0 && !mdL.empty()) {
        std::vector<MatState> ms;
        PipeOpts oc;
        oc.fixed_levels = ofix.fixed_levels;
        oc.pool_cap = ofix.pool_cap;
        oc.nwaves = ofix.nwaves_used;
        oc.hint = ofix.hint;
        oc.hint_blocks = ofix.hint_blocks;
        oc.gw = &g->gw;
        rc = pipeline_once(mdL, ms, st, oc);
    }
    if (join) {
        if (rc == 0 && hipStreamWaitEvent(st, join, 0) != hipSuccess) rc = 2;
        event_pool().put(join);
    }
    hipGraph_t graph = nullptr;
    const hipError_t ec = hipStreamEndCapture(st, &graph);
    if (rc != 0 || ec != hipSuccess || !graph) {
        if (graph) (void)hipGraphDestroy(graph);
        return 2;
    }
    g->graph = graph;
    MR3_CHECK(hipGraphInstantiateWithFlags(&g->exec, graph, 0));
    return 0;
}

// Replays a captured batch and reads its results: the states of the small route's and
// the pipeline's problems; returns 4 if the pipeline's tree needed more levels or slots
// than the capture had (its results are then not valid).
int graph_launch(BatchGraph* g, hipStream_t st, std::vector<MatState>& msS,
                 std::vector<MatState>& msL) {
    MR3_CHECK(hipGraphLaunch(g->exec, st));
    MR3_CHECK(hipStreamSynchronize(st));
    msS.assign(g->sg.hms, g->sg.hms + g->sg.nmat);
    msL.assign(g->gw.hms, g->gw.hms + g->gw.nmat);
    if (g->gw.nmat > 0) {
        if (*g->gw.hflag) return 4;
        for (int j = 0; j <= g->levels && j <= kMaxLevels; ++j)
            if (g->gw.hpin[j].err) return 2;
        if (g->levels <= kMaxLevels && g->gw.hpin[g->levels].nC > 0) return 4;
    }
    return 0;
}

BatchGraph* graph_new() { return new BatchGraph(); }
void graph_delete(BatchGraph* g) { delete g; }

void profile_enable(int on) {
    const int mode = (on == 1 || on == 2) ? on : 0;
    g_prof.on.store(mode);
    int flag = mode == 1 ? 1 : 0;
    (void)hipMemcpyToSymbol(HIP_SYMBOL(g_count_steps), &flag, sizeof(int));
}

void profile_reset() {
    std::lock_guard<std::mutex> lk(g_prof.mu);
    g_prof.acc.clear();
    for (auto& v : g_prof.steps) v = 0;
}

int profile_report(char* buf, int len) {
    static const char* kinds[kStepKinds] = {"coarse", "refine", "endref", "r1v",
                                            "r1v_bisect", "rrf", "rrj", "refine_cl",
                                            "r1v_wave", "bisect_wave", "task_ticks",
                                            "tasks", "t_bisect", "t_r1v", "t_solve", "rrf_rows",
                                            "sm_dqds", "sm_neg", "sm_sturm", "sm_r1v", "sm_rrf",
                                            "tail_r1v", "tail_bisect", "tailb_r1v", "tailb_bisect",
                                            "fallbacks"};
    std::string js = "{\"kernels\": {";
    std::lock_guard<std::mutex> lk(g_prof.mu);
    bool first = true;
    for (auto& kv : g_prof.acc) {
        char tmp[256];
        snprintf(tmp, sizeof(tmp), "%s\"%s\": {\"ms\": %.6f, \"launches\": %lld}", first ? "" : ", ",
                 kv.first.c_str(), kv.second.ms, kv.second.launches);
        js += tmp;
        first = false;
    }
    js += "}, \"steps\": {";
    for (int i = 0; i < kStepKinds; ++i) {
        char tmp[96];
        snprintf(tmp, sizeof(tmp), "%s\"%s\": %llu", i ? ", " : "", kinds[i], g_prof.steps[i]);
        js += tmp;
    }
    js += "}}";
    if ((int)js.size() + 1 > len) return -(int)(js.size() + 1);
    memcpy(buf, js.c_str(), js.size() + 1);
    return (int)js.size();
}

int set_tuning(const mr3_tuning* t) {
    if (!t) return 1;
    if (t->side_streams < 0 || t->side_streams > kSide) return 1;
    if (t->tail_singletons < 0 || t->tail_singletons > (1 << 24)) return 1;
    if (t->group_lanes < 4096 || t->rrf_group_max < 0 || t->singleton_waves < 64) return 1;
    if (t->small_first < 0 || t->small_first > (1 << 20)) return 1;
    (void)tuning_snapshot();   // environment defaults first, so that they do not override
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune = *t;
    return 0;
}

void get_tuning(mr3_tuning* t) {
    if (t) *t = tuning_snapshot();
}

}  // namespace mr3
#include "mr3_noalloc.hip"
