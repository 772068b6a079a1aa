// mr3_driver.hip -- host side of the MI355X MRRR solver: the batched device pipeline
// and the drop-in mrrr() entry point (HPAC/mr3smp SRC/mrrr.c:87-296).
//
// A solve runs the kernels of mr3_kernels.hip on one stream.  The representation tree
// is walked level by level: after the classify kernel of a level the host reads four
// counters (singletons, clusters, eigen entries, arena size) and launches the next
// level with exact grids.  That is one small device-to-host copy per tree level and
// per batch, independent of the number of matrices in the batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/mr3hip.h"
#include "mr3_internal.h"

#include "mr3_kernels.hip"

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

static inline unsigned cdiv(long long a, long long b) { return (unsigned)((a + b - 1) / b); }

// Device memory in one stream-ordered pool: buffers of a solve are taken with
// hipMallocAsync and returned with hipFreeAsync on the same stream, so that
// back-to-back solves reuse memory without device-wide synchronisation.
struct Alloc {
    hipStream_t st;
    std::vector<void*> ptrs;
    bool ok = true;
    explicit Alloc(hipStream_t s) : st(s) {}
    template <class T>
    T* get(size_t count) {
        void* p = nullptr;
        if (count == 0) count = 1;
        if (hipMallocAsync(&p, count * sizeof(T), st) != hipSuccess) { ok = false; return nullptr; }
        ptrs.push_back(p);
        return (T*)p;
    }
    void release() {
        for (void* p : ptrs) (void)hipFreeAsync(p, st);
        ptrs.clear();
    }
    ~Alloc() { release(); }
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
static size_t singleton_scratch_budget() {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) return (size_t)1 << 30;
    size_t b = fr / 4;
    const size_t cap = (size_t)32 << 30;
    return std::min(b, cap);
}

static int env_int(const char* name, int dflt) {
    const char* v = getenv(name);
    return v ? atoi(v) : dflt;
}

// Runs the whole pipeline for the problems described by md (host copy); the device
// outputs land in the caller's arrays named in md; states are copied back to ms.
int run_pipeline(std::vector<MatDesc>& md, std::vector<MatState>& ms, hipStream_t st) {
    configure_pool_once();
    const int nmat = (int)md.size();
    long long ntot = 0;
    int nmax = 0;
    bool any_v = false, any_n = false;
    for (auto& d : md) {
        d.off = (int)ntot;
        ntot += d.n;
        nmax = std::max(nmax, d.n);
        any_v |= (d.jobz == kV);
        any_n |= (d.jobz == kN);
    }
    if (ntot > 0x7fffffff / 2) return 2;
    const int N = (int)ntot;
    Alloc A(st);
    MatDesc* dmd = A.get<MatDesc>(nmat);
    MatState* dms = A.get<MatState>(nmat);
    double* D = A.get<double>(N);
    double* E = A.get<double>(N);
    double* E2 = A.get<double>(N);
    double* gers = A.get<double>(2 * (size_t)N);
    double* Dcopy = A.get<double>(N);
    double* E2copy = A.get<double>(N);
    double* DLr = A.get<double>(N);
    double* DLLr = A.get<double>(N);
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
    if (!A.ok) return 2;
    MR3_CHECK(hipMemcpyAsync(dmd, md.data(), sizeof(MatDesc) * nmat, hipMemcpyHostToDevice, st));

    hipLaunchKernelGGL(k_prep, dim3(nmat), dim3(256), 0, st, dmd, dms, D, E, E2, gers, Dcopy,
                       E2copy, isplit, blk, W, Werr, Wcoarse, iblock, Windex, smat);
    hipLaunchKernelGGL(k_coarse, dim3(cdiv(N, 256)), dim3(256), 0, st, dmd, dms, D, E2, blk, smat,
                       iblock, Windex, W, Werr, Wcoarse, N);
    hipLaunchKernelGGL(k_discard, dim3(cdiv(nmat, 64)), dim3(64), 0, st, dmd, dms, blk, W, Werr,
                       Wcoarse, iblock, Windex, nmat);
    MR3_CHECK(hipGetLastError());
    ms.resize(nmat);
    MR3_CHECK(hipMemcpyAsync(ms.data(), dms, sizeof(MatState) * nmat, hipMemcpyDeviceToHost, st));
    MR3_CHECK(hipStreamSynchronize(st));
    // block list for the root kernel: (matrix offset + block index) of every block
    std::vector<int> blist;
    int maxbl = 1;
    for (int k = 0; k < nmat; ++k) {
        if (ms[k].done || ms[k].status != kOk) continue;
        for (int j = 0; j < ms[k].nsplit; ++j) blist.push_back(md[k].off + j);
        maxbl = std::max(maxbl, ms[k].maxbl);
    }
    const int nbl = (int)blist.size();
    int* dbl = nullptr;
    if (nbl > 0) {
        dbl = A.get<int>(nbl);
        if (!A.ok) return 2;
        MR3_CHECK(hipMemcpyAsync(dbl, blist.data(), sizeof(int) * nbl, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_root, dim3(nbl), dim3(64), 0, st, dmd, dms, blk, smat, dbl, D, E, DLr,
                           DLLr, W, Werr, Wgap, nbl);
        hipLaunchKernelGGL(k_refine_root, dim3(cdiv(N, 256)), dim3(256), 0, st, dmd, dms, blk, smat,
                           iblock, Windex, D, DLLr, W, Werr, Wgap, Wtmp, Werrtmp, N);
        hipLaunchKernelGGL(k_gaps_root, dim3(cdiv(N, 256)), dim3(256), 0, st, dmd, dms, blk, smat,
                           iblock, Wtmp, Werrtmp, W, Werr, Wgap, N);
        if (any_n)
            hipLaunchKernelGGL(k_shift_n, dim3(cdiv(N, 256)), dim3(256), 0, st, dmd, dms, blk, smat,
                               iblock, W, N);
        MR3_CHECK(hipGetLastError());
    }
    if (any_v && nbl > 0) {
        hipLaunchKernelGGL(k_zindex, dim3(nmat), dim3(256), 0, st, dmd, dms, blk, W, iblock, Windex,
                           Zidx);
        // task lists: singletons (<= N), clusters (<= N/2 + 1), eigen entries (<= N)
        const int capS = N, capC = N / 2 + 1, capE = N;
        STask* Sl = A.get<STask>(capS);
        ClTask* Cl[2] = {A.get<ClTask>(capC), A.get<ClTask>(capC)};
        int* es[2] = {A.get<int>(capE), A.get<int>(capE)};
        int* ec[2] = {A.get<int>(capE), A.get<int>(capE)};
        LevelCounters* dcnt = A.get<LevelCounters>(1);
        unsigned long long* bump = A.get<unsigned long long>(1);
        if (!A.ok) return 2;
        // eigenvector scratch: 3 * maxbl doubles per lane, 64 lanes per wave
        const size_t per_wave = (size_t)3 * maxbl * 64 * sizeof(double);
        const size_t budget = singleton_scratch_budget();
        int max_waves = (int)std::max<size_t>(1, std::min<size_t>(budget / per_wave,
                                                                  (size_t)env_int("MR3_SWAVES", 8192)));
        const int want_waves = (N + 63) / 64;
        const int nwaves = std::min(max_waves, want_waves);
        double* scratch = A.get<double>((size_t)nwaves * 3 * maxbl * 64);
        if (!A.ok) return 2;
        LevelCounters hc;
        auto reset = [&]() { return hipMemsetAsync(dcnt, 0, sizeof(LevelCounters), st); };
        MR3_CHECK(reset());
        TreeOut out;
        out.S = Sl; out.C = Cl[0]; out.eig_slot = es[0]; out.eig_cl = ec[0]; out.cnt = dcnt;
        out.capS = capS; out.capC = capC; out.capEig = capE;
        hipLaunchKernelGGL(k_classify_root, dim3(cdiv(nbl, 64)), dim3(64), 0, st, dmd, dms, blk,
                           smat, dbl, gers, D, E, DLr, DLLr, W, Werr, Wgap, Wsh, out, nbl);
        MR3_CHECK(hipGetLastError());
        MR3_CHECK(hipMemcpyAsync(&hc, dcnt, sizeof(hc), hipMemcpyDeviceToHost, st));
        MR3_CHECK(hipStreamSynchronize(st));
        if (hc.err) return 2;
        int cur = 0;
        std::vector<double*> arenas;
        for (int level = 0;; ++level) {
            const int nS = hc.nS, nC = hc.nC, nE = hc.nEig;
            const long long arena_need = hc.arena;
            if (nS > 0) {
                int g = std::min(nwaves, (nS + 63) / 64);
                hipLaunchKernelGGL(k_singleton, dim3(g), dim3(64), 0, st, Sl, nS, dmd, dms, Windex,
                                   Zidx, W, Wsh, Werr, Wgap, scratch, maxbl);
                MR3_CHECK(hipGetLastError());
            }
            if (nC == 0) break;
            if (level > 64) return 2;
            double* arena = A.get<double>((size_t)arena_need);
            if (!A.ok) return 2;
            MR3_CHECK(hipMemsetAsync(bump, 0, sizeof(unsigned long long), st));
            hipLaunchKernelGGL(k_arena_assign, dim3(cdiv(nC, 256)), dim3(256), 0, st, Cl[cur], nC,
                               arena, bump);
            hipLaunchKernelGGL(k_newrrr, dim3(nC), dim3(64), 0, st, Cl[cur], nC, Windex, Wsh, Werr,
                               Wgap, dms, dmd);
            hipLaunchKernelGGL(k_refine_cl, dim3(cdiv(nE, 256)), dim3(256), 0, st, Cl[cur], es[cur],
                               ec[cur], nE, Windex, Wsh, Werr, Wgap, dms, Wtmp, Werrtmp);
            hipLaunchKernelGGL(k_gaps_cl, dim3(cdiv(nE, 256)), dim3(256), 0, st, Cl[cur], es[cur],
                               ec[cur], nE, Wtmp, Werrtmp, Wsh, Werr, Wgap, W);
            MR3_CHECK(reset());
            const int nxt = cur ^ 1;
            out.C = Cl[nxt]; out.eig_slot = es[nxt]; out.eig_cl = ec[nxt];
            hipLaunchKernelGGL(k_classify, dim3(cdiv(nC, 64)), dim3(64), 0, st, Cl[cur], nC, Wgap,
                               Wsh, out);
            MR3_CHECK(hipGetLastError());
            MR3_CHECK(hipMemcpyAsync(&hc, dcnt, sizeof(hc), hipMemcpyDeviceToHost, st));
            MR3_CHECK(hipStreamSynchronize(st));
            if (hc.err) return 2;
            cur = nxt;
        }
    }
    if (nbl > 0) {
        hipLaunchKernelGGL(k_rrj, dim3(cdiv(N, 256)), dim3(256), 0, st, dmd, dms, blk, smat, iblock,
                           Windex, Dcopy, E2copy, W, Werr, N);
        hipLaunchKernelGGL(k_finalize, dim3(nmat), dim3(256), 0, st, dmd, dms, W, Zidx, blk, iblock,
                           Windex);
        MR3_CHECK(hipGetLastError());
    }
    MR3_CHECK(hipMemcpyAsync(ms.data(), dms, sizeof(MatState) * nmat, hipMemcpyDeviceToHost, st));
    MR3_CHECK(hipStreamSynchronize(st));
    return 0;
}

}  // namespace mr3
