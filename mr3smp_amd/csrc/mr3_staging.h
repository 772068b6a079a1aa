// mr3_staging.h -- device-to-host copies of large results through pinned memory.
//
// A copy straight into pageable host memory runs at a few GB/s, most of it spent
// faulting the destination pages in on one thread.  Large results therefore go
// through a pinned double buffer: while one chunk is in flight from the device,
// several host threads copy the previous one into place.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace mr3 {

constexpr size_t kStageChunk = 64ull << 20;   // staging chunk
constexpr size_t kStageDirect = 8ull << 20;   // results up to this size are copied directly
constexpr int kCopyThreads = 8;               // host threads filling one chunk

// Pinned double buffer (one per thread that copies).
struct Staging {
    char* buf[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    bool ok = false;
    void init() {
        ok = hipHostMalloc((void**)&buf[0], kStageChunk, hipHostMallocDefault) == hipSuccess &&
             hipHostMalloc((void**)&buf[1], kStageChunk, hipHostMallocDefault) == hipSuccess &&
             hipEventCreateWithFlags(&ev[0], hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&ev[1], hipEventDisableTiming) == hipSuccess;
    }
    void release() {
        for (int k = 0; k < 2; ++k) {
            if (buf[k]) (void)hipHostFree(buf[k]);
            if (ev[k]) (void)hipEventDestroy(ev[k]);
            buf[k] = nullptr;
            ev[k] = nullptr;
        }
        ok = false;
    }
};

// A process-wide pool of staging buffers for the calls that copy large results back
// (the pinned memory is kept for the process' lifetime).  A buffer whose pinned
// allocation failed has ok false, and the copies through it go direct.
inline std::mutex& staging_mu() {
    static std::mutex m;
    return m;
}
inline std::vector<Staging*>& staging_pool() {
    static std::vector<Staging*> v;
    return v;
}
inline Staging* staging_get() {
    {
        std::lock_guard<std::mutex> lk(staging_mu());
        auto& v = staging_pool();
        if (!v.empty()) {
            Staging* s = v.back();
            v.pop_back();
            return s;
        }
    }
    Staging* s = new Staging();
    s->init();
    return s;
}
inline void staging_put(Staging* s) {
    std::lock_guard<std::mutex> lk(staging_mu());
    staging_pool().push_back(s);
}

// Runs f(t) for t = 0 .. kCopyThreads-1 on as many host threads.
template <class F>
inline void par_for(F f) {
    std::vector<std::thread> th;
    for (int t = 1; t < kCopyThreads; ++t) th.emplace_back([=] { f(t); });
    f(0);
    for (auto& x : th) x.join();
}

inline void par_copy(char* dst, const char* src, size_t bytes) {
    const size_t per = (bytes + kCopyThreads - 1) / kCopyThreads;
    par_for([=](int t) {
        const size_t o = per * t;
        if (o < bytes) memcpy(dst + o, src + o, std::min(per, bytes - o));
    });
}

// A contiguous result of `bytes` bytes, after the work queued on st; 0 on success.
inline int copy_out(void* dst, const void* src, size_t bytes, hipStream_t st, Staging& sg) {
    if (bytes <= kStageDirect || !sg.ok)
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, st) == hipSuccess ? 0 : 2;
    char* d = (char*)dst;
    const char* s = (const char*)src;
    size_t prev_o = 0, prev_c = 0;
    int k = 0;
    for (size_t o = 0; o < bytes; o += kStageChunk, ++k) {
        const size_t c = std::min(kStageChunk, bytes - o);
        const int b = k & 1;
        if (hipMemcpyAsync(sg.buf[b], s + o, c, hipMemcpyDeviceToHost, st) != hipSuccess) return 2;
        if (hipEventRecord(sg.ev[b], st) != hipSuccess) return 2;
        if (k > 0) {
            if (hipEventSynchronize(sg.ev[b ^ 1]) != hipSuccess) return 2;
            par_copy(d + prev_o, sg.buf[b ^ 1], prev_c);
        }
        prev_o = o;
        prev_c = c;
    }
    if (hipEventSynchronize(sg.ev[(k - 1) & 1]) != hipSuccess) return 2;
    par_copy(d + prev_o, sg.buf[(k - 1) & 1], prev_c);
    return 0;
}

// cols columns of rows doubles: from the device (contiguous, leading dimension rows)
// into host memory with leading dimension ld, after the work queued on st; 0 on success.
inline int copy_cols_out(double* dst, size_t ld, const double* src, size_t rows, size_t cols,
                         hipStream_t st, Staging& sg) {
    const size_t colb = rows * sizeof(double);
    const size_t bytes = colb * cols;
    if (bytes == 0) return 0;
    if (ld == rows) return copy_out(dst, src, bytes, st, sg);
    if (bytes <= kStageDirect || !sg.ok || colb > kStageChunk) {
        return hipMemcpy2DAsync(dst, ld * sizeof(double), src, colb, colb, cols,
                                hipMemcpyDeviceToHost, st) == hipSuccess ? 0 : 2;
    }
    const size_t per = kStageChunk / colb;   // columns per chunk
    auto place = [&](int b, size_t c0, size_t cc) {
        const char* from = sg.buf[b];
        par_for([=](int t) {
            for (size_t j = t; j < cc; j += kCopyThreads) memcpy(dst + (c0 + j) * ld, from + j * colb, colb);
        });
    };
    size_t prev_c0 = 0, prev_cc = 0;
    int k = 0;
    for (size_t c0 = 0; c0 < cols; c0 += per, ++k) {
        const size_t cc = std::min(per, cols - c0);
        const int b = k & 1;
        if (hipMemcpyAsync(sg.buf[b], src + c0 * rows, cc * colb, hipMemcpyDeviceToHost, st) !=
            hipSuccess)
            return 2;
        if (hipEventRecord(sg.ev[b], st) != hipSuccess) return 2;
        if (k > 0) {
            if (hipEventSynchronize(sg.ev[b ^ 1]) != hipSuccess) return 2;
            place(b ^ 1, prev_c0, prev_cc);
        }
        prev_c0 = c0;
        prev_cc = cc;
    }
    if (hipEventSynchronize(sg.ev[(k - 1) & 1]) != hipSuccess) return 2;
    place((k - 1) & 1, prev_c0, prev_cc);
    return 0;
}

}  // namespace mr3
