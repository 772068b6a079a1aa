// mr3_api.cpp -- the C-ABI entry points of libmr3hip.so (declared in include/mr3hip.h).
//
// mrrr() keeps the reference prototype and argument semantics (HPAC/mr3smp
// INCLUDE/mrrr.h:62-65, SRC/mrrr.c:87-138): it validates its arguments exactly as the
// reference does, answers the jobz="C" query for range A/I on the host, and otherwise
// copies T to the device, runs the batched pipeline on one problem and copies W, Z
// and Zsupp back.
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/mr3hip.h"
#include "mr3_args.h"
#include "mr3_internal.h"
#include "mr3_staging.h"

namespace mr3 {
int run_pipeline(std::vector<MatDesc>& md, std::vector<MatState>& ms, hipStream_t st,
                 SmallJob* small);
int small_begin(SmallJob& job, const std::vector<MatDesc>& md, hipStream_t st, bool defer);
int small_end(SmallJob& job, std::vector<MatState>& ms, hipStream_t st);
constexpr int kSmallRouteBelow = 500;   // mrrr.c:149-153 (DSTEMR_IF_SMALLER)
struct BatchGraph;
BatchGraph* graph_new();
void graph_delete(BatchGraph* g);
int graph_create(std::vector<MatDesc>& mdS, std::vector<MatDesc>& mdL, hipStream_t st,
                 BatchGraph* g);
int graph_launch(BatchGraph* g, hipStream_t st, std::vector<MatState>& msS,
                 std::vector<MatState>& msL);
void profile_enable(int on);
int set_tuning(const mr3_tuning* t);
void get_tuning(mr3_tuning* t);
void profile_reset();
int profile_report(char* buf, int len);

static hipStream_t thread_stream() {
    thread_local hipStream_t s = nullptr;
    thread_local bool tried = false;
    if (!tried) {
        tried = true;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
    }
    return s;
}

}  // namespace mr3

using namespace mr3;

extern "C" int mr3_solve_batch(mr3_problem* probs, int nprob, void* stream) {
    if (nprob <= 0) return 0;
    hipStream_t st = stream ? (hipStream_t)stream : thread_stream();
    // n < 500 takes the reference's dstemr route (mrrr.c:149-153), larger n the
    // task-parallel pipeline; both run on the device at the same time
    std::vector<MatDesc> mdS, mdL;
    std::vector<int> idxS, idxL;
    int worst = 0;
    for (int i = 0; i < nprob; ++i) {
        mr3_problem& p = probs[i];
        const int how = prepare_problem(p);
        if (how > 0) { worst = std::max(worst, 1); continue; }
        if (how < 0) continue;
        if (p.n < kSmallRouteBelow) { mdS.push_back(to_desc(p)); idxS.push_back(i); }
        else { mdL.push_back(to_desc(p)); idxL.push_back(i); }
    }
    if (mdS.empty() && mdL.empty()) return worst;
    SmallJob job;
    const int rcS = small_begin(job, mdS, st, !mdL.empty());
    std::vector<MatState> msL, msS;
    const int rcL = mdL.empty() ? 0 : run_pipeline(mdL, msL, st, rcS == 0 ? &job : nullptr);
    const int rcE = (rcS == 0) ? small_end(job, msS, st) : rcS;
    auto fill = [&](const std::vector<int>& idx, const std::vector<MatState>& ms, int rc) {
        for (size_t j = 0; j < idx.size(); ++j) {
            mr3_problem& p = probs[idx[j]];
            take_state(p, rc == 0 ? ms[j] : MatState(), rc);
            if (p.info) worst = std::max(worst, p.info);
        }
    };
    fill(idxS, msS, rcE);
    fill(idxL, msL, rcL);
    return worst;
}

// The problems of a batch split by route: host-answered and invalid ones are done here
// (their outputs set), n < 500 to the small route, the rest to the pipeline.
static int split_batch(mr3_problem* probs, int nprob, std::vector<MatDesc>& mdS,
                       std::vector<MatDesc>& mdL, std::vector<int>& idxS, std::vector<int>& idxL) {
    int worst = 0;
    for (int i = 0; i < nprob; ++i) {
        mr3_problem& p = probs[i];
        const int how = prepare_problem(p);
        if (how > 0) { worst = std::max(worst, 1); continue; }
        if (how < 0) continue;
        if (p.n < kSmallRouteBelow) { mdS.push_back(to_desc(p)); idxS.push_back(i); }
        else { mdL.push_back(to_desc(p)); idxL.push_back(i); }
    }
    return worst;
}

struct mr3_graph {
    BatchGraph* g = nullptr;
    int nprob = 0;
    std::vector<int> idxS, idxL;
};

extern "C" int mr3_graph_create(mr3_problem* probs, int nprob, void* stream, mr3_graph** out) {
    if (!out) return 1;
    *out = nullptr;
    if (nprob <= 0) return 1;
    hipStream_t st = stream ? (hipStream_t)stream : thread_stream();
    if (!st) return 2;
    mr3_graph* h = new mr3_graph();
    h->nprob = nprob;
    std::vector<MatDesc> mdS, mdL;
    int worst = split_batch(probs, nprob, mdS, mdL, h->idxS, h->idxL);
    h->g = graph_new();
    const int rc = graph_create(mdS, mdL, st, h->g);
    if (rc != 0) {
        graph_delete(h->g);
        delete h;
        for (int i = 0; i < nprob; ++i)
            if (probs[i].info == 0) probs[i].info = 2;
        return 2;
    }
    *out = h;
    // the outputs of the last solve run by the creation (the capture itself runs nothing)
    return mr3_graph_launch(h, probs, stream) == 0 ? worst : 2;
}

extern "C" int mr3_graph_launch(mr3_graph* h, mr3_problem* probs, void* stream) {
    if (!h || !probs) return 1;
    hipStream_t st = stream ? (hipStream_t)stream : thread_stream();
    if (!st) return 2;
    std::vector<MatDesc> mdS, mdL;
    std::vector<int> idxS, idxL;
    int worst = split_batch(probs, h->nprob, mdS, mdL, idxS, idxL);
    if (idxS != h->idxS || idxL != h->idxL) return 1;   // not the captured problems
    std::vector<MatState> msS, msL;
    const int rc = graph_launch(h->g, st, msS, msL);
    for (size_t j = 0; j < idxS.size(); ++j) {
        mr3_problem& p = probs[idxS[j]];
        take_state(p, rc == 2 ? MatState() : msS[j], rc == 2 ? 2 : 0);
        if (p.info) worst = std::max(worst, p.info);
    }
    for (size_t j = 0; j < idxL.size(); ++j) {
        mr3_problem& p = probs[idxL[j]];
        take_state(p, rc == 0 ? msL[j] : MatState(), rc == 0 ? 0 : 2);
        if (p.info) worst = std::max(worst, p.info);
    }
    return rc == 4 ? 4 : worst;
}

extern "C" void mr3_graph_destroy(mr3_graph* h) {
    if (!h) return;
    graph_delete(h->g);
    delete h;
}

extern "C" int mr3_solve_one(mr3_problem* prob, void* stream) {
    return mr3_solve_batch(prob, 1, stream);
}

extern "C" int mrrr(char* jobz, char* range, int* np, double* D, double* E, double* vl,
                    double* vu, int* il, int* iu, int* tryrac, int* m, double* W, double* Z,
                    int* ldz, int* Zsupp) {
    const int n = *np;
    const int jz = upper(jobz[0]), rg = upper(range[0]);
    if (check_args(jz, rg, n, rg == 'V' ? *vl : 0.0, rg == 'V' ? *vu : 0.0,
                   rg == 'I' ? *il : 1, rg == 'I' ? *iu : 1))
        return 1;
    if (jz == 'C' && rg != 'V') {   // mrrr.c:130-138
        *m = (rg == 'A') ? n : *iu - *il + 1;
        return 0;
    }
    hipStream_t st = thread_stream();
    if (!st) return 2;
    const bool wantz = (jz == 'V');
    // columns of the device Z: the index range, or for range V the count of the
    // interval plus a margin (the solve is redone with n columns if it finds more)
    int mcols = !wantz ? 0 : (rg == 'I') ? (*iu - *il + 1)
                        : (rg == 'V') ? range_v_columns(n, D, E, *vl, *vu) : n;
    double *dD = nullptr, *dE = nullptr, *dW = nullptr, *dZ = nullptr;
    int* dZs = nullptr;
    int rc = 0;
    auto fail = [&](int code) {
        if (dD) (void)hipFreeAsync(dD, st);
        if (dE) (void)hipFreeAsync(dE, st);
        if (dW) (void)hipFreeAsync(dW, st);
        if (dZ) (void)hipFreeAsync(dZ, st);
        if (dZs) (void)hipFreeAsync(dZs, st);
        (void)hipStreamSynchronize(st);
        return code;
    };
    if (hipMallocAsync((void**)&dD, sizeof(double) * n, st) != hipSuccess) return fail(2);
    if (hipMallocAsync((void**)&dE, sizeof(double) * n, st) != hipSuccess) return fail(2);
    if (hipMallocAsync((void**)&dW, sizeof(double) * n, st) != hipSuccess) return fail(2);
    if (hipMemcpyAsync(dD, D, sizeof(double) * n, hipMemcpyHostToDevice, st) != hipSuccess) return fail(2);
    if (n > 1 && hipMemcpyAsync(dE, E, sizeof(double) * (n - 1), hipMemcpyHostToDevice, st) != hipSuccess)
        return fail(2);
    mr3_problem p;
    for (int attempt = 0;; ++attempt) {
        if (wantz) {
            if (hipMallocAsync((void**)&dZ, sizeof(double) * (size_t)n * mcols, st) != hipSuccess) return fail(2);
            if (hipMallocAsync((void**)&dZs, sizeof(int) * 2 * (size_t)mcols, st) != hipSuccess) return fail(2);
        }
        memset(&p, 0, sizeof(p));
        p.n = n; p.jobz = (char)jz; p.range = (char)rg;
        p.vl = *vl; p.vu = *vu; p.il = *il; p.iu = *iu; p.tryrac = *tryrac;
        p.ldz = n; p.mcols = mcols;
        p.D = dD; p.E = dE; p.W = dW; p.Z = dZ; p.Zsupp = dZs;
        rc = mr3_solve_batch(&p, 1, st);
        if (rc == 0 || !(wantz && rg == 'V' && attempt == 0 && p.m > mcols)) break;
        (void)hipFreeAsync(dZ, st);
        (void)hipFreeAsync(dZs, st);
        dZ = nullptr;
        dZs = nullptr;
        mcols = n;
    }
    if (rc != 0) return fail(rc == 1 ? 1 : 2);
    *m = p.m;
    *vl = p.vl_out; *vu = p.vu_out; *il = p.il_out; *iu = p.iu_out; *tryrac = p.tryrac_out;
    if (jz != 'C' && p.m > 0) {
        if (hipMemcpyAsync(W, dW, sizeof(double) * p.m, hipMemcpyDeviceToHost, st) != hipSuccess) return fail(2);
        if (wantz) {
            const int lz = *ldz;
            // large blocks through pinned memory, several host threads placing them
            Staging* sg = (sizeof(double) * (size_t)n * p.m > kStageDirect) ? staging_get() : nullptr;
            Staging direct;
            const int rcz = copy_cols_out(Z, (size_t)lz, dZ, (size_t)n, (size_t)p.m, st, sg ? *sg : direct);
            if (sg) staging_put(sg);
            if (rcz != 0) return fail(2);
            if (hipMemcpyAsync(Zsupp, dZs, sizeof(int) * 2 * p.m, hipMemcpyDeviceToHost, st) != hipSuccess)
                return fail(2);
        }
    }
    return fail(0);
}

extern "C" void mrrr_(char* jobz, char* range, int* n, double* D, double* E, double* vl,
                      double* vu, int* il, int* iu, int* tryrac, int* m, double* W, double* Z,
                      int* ldz, int* Zsupp, int* info) {
    *info = mrrr(jobz, range, n, D, E, vl, vu, il, iu, tryrac, m, W, Z, ldz, Zsupp);
}

extern "C" void mr3_profile_enable(int on) { mr3::profile_enable(on); }
extern "C" void mr3_get_tuning(mr3_tuning* t) { mr3::get_tuning(t); }
extern "C" int mr3_set_tuning(const mr3_tuning* t) { return mr3::set_tuning(t); }
extern "C" void mr3_profile_reset(void) { mr3::profile_reset(); }
extern "C" int mr3_profile_report(char* buf, int len) { return mr3::profile_report(buf, len); }

extern "C" const char* mr3_version(void) { return "mr3hip 0.1 (gfx950)"; }

extern "C" double mr3_now(void) {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}
