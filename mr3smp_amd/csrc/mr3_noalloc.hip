// mr3_noalloc.hip -- the small-matrix route (k_small, n < 500) run without allocating:
// every device array it needs lives in one static device buffer, `scratch`, that each
// GPU the library runs on holds until the process ends.
//
// Layout of scratch for a call with nm problems (256-byte aligned parts):
//   MatDesc[nm] | MatState[nm] | int list[nm] | long long slice offset[nm] | slices,
// one slice per problem (mr3sm::carve lays it out; k_small writes every entry it reads,
// so nothing has to start at zero).  Calls that overlap on one GPU share the buffer
// and race.  Included into mr3_driver.hip's translation unit, beside k_small.
#pragma once
#include <hip/hip_runtime.h>

#include <mutex>

#include "mr3_args.h"
#include "mr3_internal.h"

namespace mr3 {

constexpr long long kNoAllocScratchBytes = 64LL << 20;   // per GPU
constexpr int kNoAllocMaxProblems = 1024;
constexpr int kNoAllocMaxDevices = 64;

static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

// The device address of each GPU's copy of scratch, taken once per device with that
// device current.  The name `scratch` in host code is only a stand-in for the symbol.
static std::mutex g_noalloc_mu;
static char* g_noalloc_ptr[kNoAllocMaxDevices];

static int noalloc_scratch(char** out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kNoAllocMaxDevices) return 2;
    if (!g_noalloc_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return 2;
        g_noalloc_ptr[dev] = (char*)p;
    }
    *out = g_noalloc_ptr[dev];
    return 0;
}

// Host staging of the descriptors and states (static: the entry point allocates no
// host memory either); g_noalloc_mu serialises the calls of the process.
static MatDesc g_noalloc_md[kNoAllocMaxProblems];
static MatState g_noalloc_ms[kNoAllocMaxProblems];
static int g_noalloc_idx[kNoAllocMaxProblems];
static int g_noalloc_list[kNoAllocMaxProblems];
static long long g_noalloc_off[kNoAllocMaxProblems];

// Checks and solves probs[0..nprob) on st, waiting for the results.  Returns as
// mr3_solve_batch, or 3 (nothing run, every info 3) when a problem has n >= 500, there
// are more than kNoAllocMaxProblems, or their work arrays do not fit scratch.
static int small_noalloc(mr3_problem* probs, int nprob, hipStream_t st) {
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    auto reject = [&]() {
        for (int i = 0; i < nprob; ++i) { probs[i].m = 0; probs[i].info = 3; }
        return 3;
    };
    if (nprob <= 0) return 0;
    if (nprob > kNoAllocMaxProblems) return reject();
    for (int i = 0; i < nprob; ++i)
        if (probs[i].n >= kSmallBelow) return reject();
    int worst = 0, nm = 0;
    for (int i = 0; i < nprob; ++i) {
        mr3_problem& p = probs[i];
        const int how = prepare_problem(p);
        if (how > 0) { worst = std::max(worst, 1); continue; }
        if (how < 0) continue;
        g_noalloc_md[nm] = to_desc(p);
        g_noalloc_idx[nm] = i;
        ++nm;
    }
    if (nm == 0) return worst;
    const MatDesc* md = g_noalloc_md;
    // largest first: the longest solves start first
    int* list = g_noalloc_list;
    for (int i = 0; i < nm; ++i) {
        int j = i;
        while (j > 0 && md[list[j - 1]].n < md[i].n) { list[j] = list[j - 1]; --j; }
        list[j] = i;
    }
    auto al = [](long long b) { return (b + 255) & ~255LL; };
    const long long bmd = al(sizeof(MatDesc) * nm), bms = al(sizeof(MatState) * nm);
    const long long bl = al(sizeof(int) * nm), bo = al(sizeof(long long) * nm);
    long long tot = 0;
    int nmax = 1;
    for (int j = 0; j < nm; ++j) {
        const int n = md[list[j]].n;
        g_noalloc_off[j] = tot;
        tot += small_slice_bytes(n);
        nmax = std::max(nmax, n);
    }
    if (bmd + bms + bl + bo + tot > kNoAllocScratchBytes) return reject();
    char* base = nullptr;
    int rc = noalloc_scratch(&base);
    if (rc == 0) {
        MatDesc* dmd = (MatDesc*)base;
        MatState* dms = (MatState*)(base + bmd);
        int* dlist = (int*)(base + bmd + bms);
        long long* doff = (long long*)(base + bmd + bms + bl);
        char* slices = base + bmd + bms + bl + bo;
        rc = [&]() {
            MR3_CHECK(hipMemcpyAsync(dmd, md, sizeof(MatDesc) * nm, hipMemcpyHostToDevice, st));
            MR3_CHECK(hipMemcpyAsync(dlist, list, sizeof(int) * nm, hipMemcpyHostToDevice, st));
            MR3_CHECK(hipMemcpyAsync(doff, g_noalloc_off, sizeof(long long) * nm,
                                     hipMemcpyHostToDevice, st));
            const size_t lds = sizeof(double) * (size_t)mr3sm::lds_doubles(nmax);
            hipLaunchKernelGGL(k_small, dim3(nm), dim3(kSmallNT), lds, st, dmd, dms, dlist, doff,
                               slices);
            MR3_CHECK(hipGetLastError());
            MR3_CHECK(hipMemcpyAsync(g_noalloc_ms, dms, sizeof(MatState) * nm,
                                     hipMemcpyDeviceToHost, st));
            MR3_CHECK(hipStreamSynchronize(st));
            return 0;
        }();
    }
    for (int j = 0; j < nm; ++j) {
        mr3_problem& p = probs[g_noalloc_idx[j]];
        take_state(p, g_noalloc_ms[j], rc);
        if (p.info) worst = std::max(worst, p.info);
    }
    return worst;
}

}  // namespace mr3

// mr3_solve_batch for problems of order n < 500 without allocating (declared in
// include/mr3hip.h).  stream NULL: the calling thread's default stream.
extern "C" int mr3_solve_small_noalloc(mr3_problem* probs, int nprob, void* stream) {
    return mr3::small_noalloc(probs, nprob, stream ? (hipStream_t)stream : hipStreamPerThread);
}
