// mr3_serve.cpp -- request loop with continuous batching for the MRRR solver.
//
// A scheduler thread admits queued requests in arrival order whenever a concurrency
// slot is free and the device-memory budget (8 n m bytes for Z plus a per-row
// workspace estimate) has room; each admitted group runs as one batched solve on a
// worker's stream, so groups that were admitted at different times overlap on the
// device.  Requests arriving while groups run wait only for a slot, not for the
// whole queue to drain.  Results of a problem do not depend on the batch it ran
// in: every decision of the pipeline is taken per matrix.
//
// The workers are persistent threads, one per concurrency slot, each with its own
// stream (and the solver's per-thread side streams); they run a small solve when the
// server starts, so the first requests do not pay for stream creation, device
// workspace or code loading.  Large results go back to the caller's host arrays
// through a pinned double buffer: while one chunk is in flight from the device,
// several host threads copy the previous one into place (a copy straight into
// pageable memory runs at a few GB/s, most of it spent faulting the destination
// pages in on one thread).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/mr3hip.h"
#include "mr3_args.h"
#include "mr3_staging.h"

namespace mr3 {
void release_thread_streams();   // mr3_driver.hip: the solver's per-thread side streams
}

namespace {

using mr3::copy_out;
using mr3::Staging;

constexpr long long kRowBytes = 512;       // per-row workspace estimate of the pipeline

// Eigenvector columns a request takes on the device: the caller's mcols if given,
// else iu - il + 1 for range I, the interval's count plus a margin for range V (the
// reference's jobz = 'C' query, mrrr.c:130-138), n for range A.
int device_columns(const mr3_request& r) {
    if (!(r.jobz == 'V' || r.jobz == 'v')) return 0;
    if (r.mcols > 0) return r.mcols;
    const int rg = mr3::upper(r.range);
    if (rg == 'I') return std::max(1, r.iu - r.il + 1);
    if (rg == 'V' && r.n > 0 && r.vu > r.vl) return mr3::range_v_columns(r.n, r.D, r.E, r.vl, r.vu);
    return std::max(1, r.n);
}

long long footprint(const mr3_request& r, int cols) {
    return 8LL * r.n * cols + kRowBytes * r.n + 64LL * r.n;
}

struct Entry {
    long long id;
    mr3_request* req;
    long long bytes;
    int cols;              // device eigenvector columns
};

}  // namespace

struct mr3_server {
    long long budget = 0, in_use = 0;
    int max_inflight = 2, max_batch = 256, inflight = 0, ready = 0;
    bool stop = false;
    long long next_id = 0;
    std::deque<Entry> pending;
    std::deque<std::vector<Entry>> groups;               // admitted, waiting for a worker
    std::unordered_map<long long, mr3_request*> live;   // submitted, not yet complete
    std::unordered_map<long long, mr3_request*> reqs;   // submitted, not yet waited for
    std::mutex mu;
    std::condition_variable cv_sched, cv_done, cv_work;
    std::thread sched;
    std::vector<std::thread> workers;
    int device = 0;

    void run_group(std::vector<Entry>& grp, hipStream_t st, Staging& sg);
    void scheduler();
    void worker();
};

void mr3_server::run_group(std::vector<Entry>& grp, hipStream_t st, Staging& sg) {
    const double t0 = mr3_now();
    for (auto& e : grp) e.req->t_start = t0;
    // one device allocation for the whole group
    size_t bytes = 0;
    std::vector<size_t> offs(grp.size() * 5);
    auto take = [&](size_t b) { size_t o = bytes; bytes += (b + 255) & ~(size_t)255; return o; };
    for (size_t i = 0; i < grp.size(); ++i) {
        const mr3_request& r = *grp[i].req;
        const int mc = grp[i].cols;
        offs[5 * i + 0] = take(sizeof(double) * r.n);
        offs[5 * i + 1] = take(sizeof(double) * r.n);
        offs[5 * i + 2] = take(sizeof(double) * r.n);
        offs[5 * i + 3] = take(sizeof(double) * (size_t)r.n * mc);
        offs[5 * i + 4] = take(sizeof(int) * 2 * (size_t)(mc > 0 ? mc : 1));
    }
    // transport: the group's allocation, copies and synchronisation; a failure there
    // fails every request of the group, a solver failure only its own request
    bool transport_ok = true;
    char* base = nullptr;
    if (hipMallocAsync((void**)&base, bytes ? bytes : 1, st) != hipSuccess) {
        base = nullptr;
        transport_ok = false;
    }
    std::vector<mr3_problem> probs(grp.size());
    if (transport_ok) {
        for (size_t i = 0; i < grp.size() && transport_ok; ++i) {
            const mr3_request& r = *grp[i].req;
            const bool wz = grp[i].cols > 0;
            mr3_problem& p = probs[i];
            memset(&p, 0, sizeof(p));
            p.n = r.n; p.jobz = r.jobz; p.range = r.range;
            p.vl = r.vl; p.vu = r.vu; p.il = r.il; p.iu = r.iu; p.tryrac = r.tryrac;
            p.ldz = r.n;
            p.mcols = grp[i].cols;
            p.D = (const double*)(base + offs[5 * i + 0]);
            p.E = (const double*)(base + offs[5 * i + 1]);
            p.W = (double*)(base + offs[5 * i + 2]);
            p.Z = wz ? (double*)(base + offs[5 * i + 3]) : nullptr;
            p.Zsupp = wz ? (int*)(base + offs[5 * i + 4]) : nullptr;
            if (hipMemcpyAsync((void*)p.D, r.D, sizeof(double) * r.n, hipMemcpyHostToDevice, st) != hipSuccess)
                transport_ok = false;
            if (r.n > 1 && hipMemcpyAsync((void*)p.E, r.E, sizeof(double) * (r.n - 1), hipMemcpyHostToDevice,
                                          st) != hipSuccess)
                transport_ok = false;
        }
    }
    if (transport_ok) {
        (void)mr3_solve_batch(probs.data(), (int)probs.size(), st);
        for (size_t i = 0; i < grp.size(); ++i) {
            mr3_request& r = *grp[i].req;
            mr3_problem& p = probs[i];
            // a range-V request without mcols (its host Z holds n columns) whose
            // interval held more than its count's margin runs again alone with n
            // columns (mrrr() does the same); the retry's memory is outside the budget
            char* big = nullptr;
            if (p.info == 2 && p.Z && p.m > p.mcols && r.mcols <= 0 && mr3::upper(r.range) == 'V' &&
                hipMallocAsync((void**)&big, sizeof(double) * (size_t)r.n * r.n + sizeof(int) * 2 * r.n + 256,
                               st) == hipSuccess) {
                p.mcols = r.n;
                p.Z = (double*)big;
                p.Zsupp = (int*)(big + (((sizeof(double) * (size_t)r.n * r.n) + 255) & ~(size_t)255));
                (void)mr3_solve_batch(&p, 1, st);
            }
            r.m = p.m;
            r.info = p.info;
            if (p.info == 0 && p.m > 0) {
                int c = 0;
                if (r.W) c |= copy_out(r.W, p.W, sizeof(double) * p.m, st, sg);
                if (p.Z && r.Z) c |= copy_out(r.Z, p.Z, sizeof(double) * (size_t)r.n * p.m, st, sg);
                if (p.Zsupp && r.Zsupp) c |= copy_out(r.Zsupp, p.Zsupp, sizeof(int) * 2 * p.m, st, sg);
                if (c) r.info = 2;
            }
            if (big) (void)hipFreeAsync(big, st);
        }
    }
    if (base) (void)hipFreeAsync(base, st);
    if (hipStreamSynchronize(st) != hipSuccess) transport_ok = false;
    const double t1 = mr3_now();
    std::lock_guard<std::mutex> lk(mu);
    for (auto& e : grp) {
        if (!transport_ok) e.req->info = 2;
        e.req->t_done = t1;
        in_use -= e.bytes;
        live.erase(e.id);
    }
    inflight -= 1;
    cv_done.notify_all();
    cv_sched.notify_all();
}

void mr3_server::worker() {
    (void)hipSetDevice(device);
    hipStream_t st = nullptr;
    (void)hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    Staging sg;
    sg.init();
    {
        // warm-up: a 3x3 solve creates this thread's streams and device workspace
        double h[9] = {2.0, 2.0, 2.0, -1.0, -1.0, 0.0};
        double* d = nullptr;
        if (st && hipMallocAsync((void**)&d, sizeof(double) * 32, st) == hipSuccess) {
            (void)hipMemcpyAsync(d, h, sizeof(double) * 6, hipMemcpyHostToDevice, st);
            mr3_problem p;
            memset(&p, 0, sizeof(p));
            p.n = 3; p.jobz = 'V'; p.range = 'A'; p.tryrac = 1; p.ldz = 3; p.mcols = 3;
            p.D = d; p.E = d + 3; p.W = d + 8; p.Z = d + 12; p.Zsupp = (int*)(d + 24);
            (void)mr3_solve_batch(&p, 1, st);
            (void)hipFreeAsync(d, st);
            (void)hipStreamSynchronize(st);
        }
    }
    std::unique_lock<std::mutex> lk(mu);
    ++ready;
    cv_done.notify_all();
    for (;;) {
        cv_work.wait(lk, [&] { return stop || !groups.empty(); });
        if (groups.empty()) break;   // stop, nothing left
        std::vector<Entry> grp = std::move(groups.front());
        groups.pop_front();
        lk.unlock();
        run_group(grp, st, sg);
        lk.lock();
    }
    lk.unlock();
    sg.release();
    mr3::release_thread_streams();
    if (st) (void)hipStreamDestroy(st);
}

void mr3_server::scheduler() {
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        cv_sched.wait(lk, [&] {
            if (stop) return true;
            if (pending.empty() || inflight >= max_inflight) return false;
            return in_use + pending.front().bytes <= budget;
        });
        if (stop) break;
        // admit in arrival order while the budget and the group size allow
        std::vector<Entry> grp;
        while (!pending.empty() && (int)grp.size() < max_batch &&
               in_use + pending.front().bytes <= budget) {
            grp.push_back(pending.front());
            in_use += pending.front().bytes;
            pending.pop_front();
        }
        if (grp.empty()) continue;
        inflight += 1;
        groups.push_back(std::move(grp));
        cv_work.notify_one();
    }
}

extern "C" mr3_server* mr3_server_create(long long budget_bytes, int max_inflight, int max_batch) {
    mr3_server* s = new mr3_server();
    s->budget = budget_bytes > 0 ? budget_bytes : (200LL << 30);
    s->max_inflight = max_inflight > 0 ? max_inflight : 2;
    s->max_batch = max_batch > 0 ? max_batch : 256;
    (void)hipGetDevice(&s->device);
    for (int w = 0; w < s->max_inflight; ++w) s->workers.emplace_back(&mr3_server::worker, s);
    {
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.wait(lk, [&] { return s->ready == s->max_inflight; });
    }
    s->sched = std::thread(&mr3_server::scheduler, s);
    return s;
}

extern "C" long long mr3_server_submit(mr3_server* s, mr3_request* req) {
    Entry e;
    e.req = req;
    e.cols = device_columns(*req);
    e.bytes = footprint(*req, e.cols);
    req->t_submit = mr3_now();
    req->t_start = req->t_done = 0.0;
    req->m = 0;
    req->info = -1;
    std::lock_guard<std::mutex> lk(s->mu);
    if (e.bytes > s->budget) return -1;
    e.id = s->next_id++;
    s->pending.push_back(e);
    s->live[e.id] = req;
    s->reqs[e.id] = req;
    s->cv_sched.notify_all();
    return e.id;
}

extern "C" int mr3_server_wait(mr3_server* s, long long id) {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end()) return -1;   // unknown, rejected or already waited for
    mr3_request* r = it->second;
    s->cv_done.wait(lk, [&] { return s->live.find(id) == s->live.end(); });
    s->reqs.erase(id);
    return r->info;
}

extern "C" void mr3_server_drain(mr3_server* s) {
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->live.empty(); });
}

extern "C" void mr3_server_destroy(mr3_server* s) {
    mr3_server_drain(s);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_sched.notify_all();
        s->cv_work.notify_all();
    }
    s->sched.join();
    for (auto& w : s->workers) w.join();
    delete s;
}
