// CPU build of the small-matrix route of the GPU library (mr3smp_amd/csrc/
// mr3_small_core.h), run by one thread, for the tests: it lets the device algorithm be
// compared bit for bit with the oracle without a GPU.  Test infrastructure only.
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define SM_HD
#define SM_SYNC() ((void)0)
#include "../mr3smp_amd/csrc/mr3_small_core.h"

using namespace mr3sm;

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
    Work w;
    double* lds = (double*)calloc((size_t)lds_doubles(n), sizeof(double));
    char* slice = (char*)calloc((size_t)slice_bytes(n, 1), 1);
    carve(w, lds, slice, n, 1);
    Shared sh;
    memset(&sh, 0, sizeof(sh));
    Result r = dstemr_team(0, 1, sh, p, w);
    if (getenv("SMALL_ROUTE_PASSES")) fprintf(stderr, "n=%d passes=%d\n", n, sh.passes);
    free(lds);
    free(slice);
    *m = r.m;
    if (!p.cnt) *tryrac = r.tryrac;
    return r.info ? 2 : 0;
}
