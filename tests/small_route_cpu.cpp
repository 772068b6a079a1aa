// CPU build of the small-matrix route of the GPU library (mr3smp_amd/csrc/
// mr3_small_core.h) for the tests: it lets the device algorithm be compared bit for bit
// with the oracle without a GPU.  By default one thread runs every phase in order;
// small_route_set_team(nt) makes nt POSIX threads run it as the device's team does, with
// a barrier for each of the team's synchronisations, so that the work split and the
// shared state of a 64-lane wave are exercised too.  Test infrastructure only.
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

// The team's barrier: every thread must reach it from the same source line (the core's
// synchronisations are meant to be met by the whole team at once); a thread waiting
// longer than 20 s, or threads meeting at different lines, end the process with a
// report of where each thread stands.
static int g_team = 1;
static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static pthread_cond_t g_cv = PTHREAD_COND_INITIALIZER;
static int g_arrived = 0;
static long g_gen = 0;
static int g_site[1024];
static __thread int t_tid = 0;

static void team_report(const char* what) {
    fprintf(stderr, "team barrier: %s; sites:", what);
    for (int t = 0; t < g_team; ++t) fprintf(stderr, " %d", g_site[t]);
    fprintf(stderr, "\n");
    abort();
}

static inline void team_sync(int line) {
    if (g_team <= 1) return;
    pthread_mutex_lock(&g_mu);
    g_site[t_tid] = line;
    const long gen = g_gen;
    if (++g_arrived == g_team) {
        for (int t = 1; t < g_team; ++t)
            if (g_site[t] != g_site[0]) team_report("threads meet at different lines");
        g_arrived = 0;
        ++g_gen;
        pthread_cond_broadcast(&g_cv);
    } else {
        struct timespec ts;
        clock_gettime(CLOCK_REALTIME, &ts);
        ts.tv_sec += 20;
        while (g_gen == gen)
            if (pthread_cond_timedwait(&g_cv, &g_mu, &ts) != 0 && g_gen == gen) team_report("timeout");
    }
    pthread_mutex_unlock(&g_mu);
}

#define SM_HD
#define SM_SYNC() team_sync(__LINE__)
#include "../mr3smp_amd/csrc/mr3_small_core.h"

using namespace mr3sm;

struct TeamArg {
    int tid, nt;
    Shared* sh;
    const Prob* p;
    const Work* w;
    Result r;
};

// the stage a team thread runs: 0 the whole solve, 1 the first stage of the staged
// form, 2 its second stage
static int g_stage = 0;
static int g_staged = 0;

static void* team_main(void* a) {
    TeamArg* t = (TeamArg*)a;
    t_tid = t->tid;
    if (g_stage == 0) {
        t->r = dstemr_team(t->tid, t->nt, *t->sh, *t->p, *t->w);
    } else if (g_stage == 1) {
        if (dstemr_stage_a(t->tid, t->nt, *t->sh, *t->p, *t->w, true, t->r, *t->w->STG))
            t->w->STG->fin = 1;
        if (t->w->STG->fin) t->w->STG->res = t->r;
    } else {
        t->r = dstemr_stage_b(t->tid, t->nt, *t->sh, *t->p, *t->w, t->w->STG->res, *t->w->STG);
    }
    return nullptr;
}

// runs one stage on the team (g_team threads, or the calling thread alone)
static Result run_team(int stage, Shared& sh, const Prob& p, const Work& w) {
    const int nt = g_team;
    g_stage = stage;
    Result r;
    if (nt == 1) {
        TeamArg a;
        a.tid = 0; a.nt = 1; a.sh = &sh; a.p = &p; a.w = &w;
        team_main(&a);
        r = a.r;
    } else {
        g_arrived = 0;
        TeamArg* args = new TeamArg[nt];
        pthread_t* th = new pthread_t[nt];
        for (int t = 0; t < nt; ++t) {
            args[t].tid = t; args[t].nt = nt; args[t].sh = &sh; args[t].p = &p; args[t].w = &w;
            pthread_create(&th[t], nullptr, team_main, &args[t]);
        }
        for (int t = 0; t < nt; ++t) pthread_join(th[t], nullptr);
        r = args[0].r;
        delete[] th;
        delete[] args;
    }
    return r;
}

extern "C" void small_route_set_team(int nt) { g_team = nt < 1 ? 1 : (nt > 1024 ? 1024 : nt); }
// staged != 0: run the two-stage form with every deferred dqds between the stages
extern "C" void small_route_set_staged(int staged) { g_staged = staged; }

// Same arguments as mrrr() for the matrices it sends to dstemr (n < 500).
extern "C" int small_route_mrrr(char* jobz, char* range, int* np, double* D, double* E,
                                double* vl, double* vu, int* il, int* iu, int* tryrac,
                                int* m, double* W, double* Z, int* ldz, int* Zsupp) {
    const int n = *np;
    const char jz = jobz[0] & ~0x20, rg = range[0] & ~0x20;
    Prob p;
    memset(&p, 0, sizeof(p));
    p.n = n;
    p.wantz = (jz == 'V');
    p.cnt = (jz == 'C');
    p.range = rg == 'A' ? kAll : (rg == 'V' ? kVal : kInd);
    p.il = *il; p.iu = *iu; p.vl = *vl; p.vu = *vu;
    p.tryrac = *tryrac;
    p.ldz = *ldz;
    p.mcols = n;
    p.Din = D; p.Ein = E; p.W = W; p.Z = Z; p.Zsupp = Zsupp;
    const int nt = g_team;
    const int wks = n < nt ? n : nt;   // threads whose dlar1v / dlarrf work is kept
    Work w;
    // the device's LDS and work slice start with whatever they held: poison them here
    // (a byte pattern that reads as NaN doubles and negative ints), so that a read of an
    // entry before its first write changes the results
    const char* pz = getenv("SMALL_ROUTE_POISON");
    const int pat = pz ? (int)strtol(pz, nullptr, 0) : 0xA7;
    double* lds = (double*)malloc((size_t)lds_doubles(n) * sizeof(double));
    char* slice = (char*)malloc((size_t)slice_bytes(n, wks));
    memset(lds, pat, (size_t)lds_doubles(n) * sizeof(double));
    memset(slice, pat, (size_t)slice_bytes(n, wks));
    carve(w, lds, slice, n, wks);
    Shared sh;
    memset(&sh, 0, sizeof(sh));
    Result r;
    if (!g_staged) {
        r = run_team(0, sh, p, w);
    } else {
        // the staged form as the device runs it: the first stage, D and E kept, the
        // team's LDS and shared record lost (poisoned), every deferred block's dqds on
        // its own, then the second stage with D and E back
        run_team(1, sh, p, w);
        if (w.STG->fin) {
            r = w.STG->res;
        } else {
            memcpy(w.DSAVE, w.D, sizeof(double) * n);
            memcpy(w.DSAVE + n, w.E, sizeof(double) * n);
            memset(lds, pat, (size_t)lds_doubles(n) * sizeof(double));
            memset(&sh, pat, sizeof(sh));
            for (int q = 0; q < w.STG->ndq; ++q) {
                DqRec& rec = w.DQR[q];
                rec.info = dlasq2(rec.in, w.QD + 4 * (rec.ib - 1));
            }
            memcpy(w.D, w.DSAVE, sizeof(double) * n);
            memcpy(w.E, w.DSAVE + n, sizeof(double) * n);
            r = run_team(2, sh, p, w);
        }
    }
    if (getenv("SMALL_ROUTE_PASSES")) fprintf(stderr, "n=%d passes=%d\n", n, sh.passes);
    free(lds);
    free(slice);
    *m = r.m;
    if (!p.cnt) *tryrac = r.tryrac;
    return r.info ? 2 : 0;
}
