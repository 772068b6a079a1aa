/*
 * mr3_oracle.c -- CPU oracle for the MRRR tridiagonal eigensolver.
 *
 * TEST INFRASTRUCTURE ONLY: a sequential C restatement of the reference's algorithm
 * (HPAC/mr3smp), used as the checker by tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py.  The product (mr3smp_amd/) never links or calls it.
 * Every function cites the reference file:line it restates; the scope and the
 * deliberate differences from the reference are listed in mr3_oracle.h.
 *
 * Compiled with -ffp-contract=off so that no multiply-add is fused, as in the
 * reference's x86-64 build.
 */
#include "mr3_oracle.h"
#include "mr3_oracle_int.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Singletons that fell back to bisection (process_singleton.c:186-204) since the last
 * oracle_fallbacks(1) call: a statistic that tests compare with the device's count. */
static long long o_fallbacks = 0;
long long oracle_fallbacks(int reset)
{
    const long long v = o_fallbacks;
    if (reset) o_fallbacks = 0;
    return v;
}

#define EPS     DBL_EPSILON      /* odmch.f 'P' = EPSILON(0d0)            */
#define SAFMIN  DBL_MIN          /* odmch.f 'S' and 'U' = TINY(0d0)       */

/* tunables of the reference (INCLUDE/mrrr.h:152-200, mrrr_val.c:95-99) */
#define MAX_GROWTH      64.0
#define MIN_RELGAP      1e-3
#define MAX_TRY_ROOT    6
#define STASK_NOENQUEUE 10
#define FOURTH          0.25
#define HUNDRED         100.0
#define FUDGE_ROOT      2.0
#define RAND_FACTOR     8.0
#define MAXITER_S       10       /* process_singleton.c:57 */

static inline int    imin(int a, int b) { return a < b ? a : b; }
static inline int    imax(int a, int b) { return a > b ? a : b; }

/* ------------------------------------------------------------------------------
 * dlaneg: Sturm count of L D L^T - sigma I via a twisted factorisation at r.
 * odneg.f:97-164 (blocks of BLKLEN=128 with a NaN-safe re-run, odneg.f:82,111-126).
 * ---------------------------------------------------------------------------- */
int oracle_dlaneg(int n, const double *d, const double *lld, double sigma,
                  double pivmin, int r)
{
    (void)pivmin;
    const int BLKLEN = 128;
    int negcnt = 0;
    double t = -sigma;
    for (int bj = 1; bj <= r - 1; bj += BLKLEN) {
        int neg1 = 0;
        double bsav = t;
        int jend = imin(bj + BLKLEN - 1, r - 1);
        for (int j = bj; j <= jend; j++) {
            double dplus = d[j - 1] + t;
            if (dplus < 0.0) neg1++;
            double tmp = t / dplus;
            t = tmp * lld[j - 1] - sigma;
        }
        if (isnan(t)) {
            neg1 = 0;
            t = bsav;
            for (int j = bj; j <= jend; j++) {
                double dplus = d[j - 1] + t;
                if (dplus < 0.0) neg1++;
                double tmp = t / dplus;
                if (isnan(tmp)) tmp = 1.0;
                t = tmp * lld[j - 1] - sigma;
            }
        }
        negcnt += neg1;
    }
    double p = d[n - 1] - sigma;
    for (int bj = n - 1; bj >= r; bj -= BLKLEN) {
        int neg2 = 0;
        double bsav = p;
        int jend = imax(bj - BLKLEN + 1, r);
        for (int j = bj; j >= jend; j--) {
            double dminus = lld[j - 1] + p;
            if (dminus < 0.0) neg2++;
            double tmp = p / dminus;
            p = tmp * d[j - 1] - sigma;
        }
        if (isnan(p)) {
            neg2 = 0;
            p = bsav;
            for (int j = bj; j >= jend; j--) {
                double dminus = lld[j - 1] + p;
                if (dminus < 0.0) neg2++;
                double tmp = p / dminus;
                if (isnan(tmp)) tmp = 1.0;
                p = tmp * d[j - 1] - sigma;
            }
        }
        negcnt += neg2;
    }
    double gamma = (t + sigma) + p;
    if (gamma < 0.0) negcnt++;
    return negcnt;
}

/* Sturm count on T at x, dlaebz IJOB=1 convention (odebz.f:265-277): pivots with
 * |t| < pivmin are replaced by -pivmin; a pivot <= 0 counts. */
int oracle_sturm_count_T(int n, const double *d, const double *e2, double x,
                         double pivmin)
{
    double tmp = d[0] - x;
    if (fabs(tmp) < pivmin) tmp = -pivmin;
    int cnt = (tmp <= 0.0) ? 1 : 0;
    for (int j = 2; j <= n; j++) {
        tmp = d[j - 1] - e2[j - 2] / tmp - x;
        if (fabs(tmp) < pivmin) tmp = -pivmin;
        if (tmp <= 0.0) cnt++;
    }
    return cnt;
}

/* ------------------------------------------------------------------------------
 * dlaruv / dlarnv (IDIST=2): 48-bit multiplicative congruential generator,
 * x_i = seed * a^i mod 2^48 with a = 33952834046453, returned as x / 2^48
 * (odruv.f:330-387); dlarnv maps u -> 2u - 1 in chunks of LV/2 = 64 (odrnv.f:60-116).
 * The multiplier table MM(i,1:4) of odruv.f holds a^i mod 2^48 in 12-bit digits;
 * here the powers are formed directly.  The X == 1.0 retry of odruv.f:362-375 can
 * not trigger: a 48-bit integer divided by 2^48 is exact in binary64 and < 1.
 * ---------------------------------------------------------------------------- */
static const unsigned long long MR3_LCG_A = 33952834046453ULL;
#define MASK48 ((1ULL << 48) - 1ULL)

void oracle_dlarnv2(int iseed[4], int n, double *x)
{
    unsigned long long seed = ((unsigned long long)iseed[0] << 36) |
                              ((unsigned long long)iseed[1] << 24) |
                              ((unsigned long long)iseed[2] << 12) |
                              (unsigned long long)iseed[3];
    const double R = 1.0 / 4096.0;
    for (int iv = 0; iv < n; iv += 64) {
        int il = imin(64, n - iv);
        unsigned long long apow = 1ULL, xi = seed;
        for (int i = 0; i < il; i++) {
            apow = (apow * MR3_LCG_A) & MASK48;
            xi = (seed * apow) & MASK48;
            double it1 = (double)((xi >> 36) & 4095ULL);
            double it2 = (double)((xi >> 24) & 4095ULL);
            double it3 = (double)((xi >> 12) & 4095ULL);
            double it4 = (double)(xi & 4095ULL);
            double u = R * (it1 + R * (it2 + R * (it3 + R * it4)));
            x[iv + i] = 2.0 * u - 1.0;
        }
        seed = xi;
    }
    iseed[0] = (int)((seed >> 36) & 4095ULL);
    iseed[1] = (int)((seed >> 24) & 4095ULL);
    iseed[2] = (int)((seed >> 12) & 4095ULL);
    iseed[3] = (int)(seed & 4095ULL);
}

/* ------------------------------------------------------------------------------
 * dlarrb: bisection refinement of eigenvalues IFIRST..ILAST of L D L^T,
 * odrrb.f:135-299 (linked list of unconverged intervals, MAXITR cap).
 * w, wgap, werr are indexed II = I - OFFSET (1-based).
 * ---------------------------------------------------------------------------- */
void oracle_dlarrb(int n, const double *d, const double *lld, int ifirst,
                   int ilast, double rtol1, double rtol2, int offset,
                   double *w, double *wgap, double *werr, double pivmin,
                   double spdiam, int twist)
{
#define W_(ii)    w[(ii) - 1]
#define WGAP_(ii) wgap[(ii) - 1]
#define WERR_(ii) werr[(ii) - 1]
    double *work = (double *)malloc(sizeof(double) * (size_t)(2 * ilast + 2));
    int *iwork = (int *)malloc(sizeof(int) * (size_t)(2 * ilast + 2));
#define WORK_(k)  work[(k)]
#define IWORK_(k) iwork[(k)]
    int maxitr = (int)((log(spdiam + pivmin) - log(pivmin)) / log(2.0)) + 2;
    double mnwdth = 2.0 * pivmin;
    int r = twist;
    if (r < 1 || r > n) r = n;

    int i1 = ifirst, nint = 0, prev = 0;
    double rgap = WGAP_(i1 - offset), lgap, gap, left, right, width, tmp, cvrgd, back;
    for (int i = i1; i <= ilast; i++) {
        int k = 2 * i, ii = i - offset;
        left = W_(ii) - WERR_(ii);
        right = W_(ii) + WERR_(ii);
        lgap = rgap;
        rgap = WGAP_(ii);
        gap = fmin(lgap, rgap);
        back = WERR_(ii);
        int negcnt;
        for (;;) {
            negcnt = oracle_dlaneg(n, d, lld, left, pivmin, r);
            if (negcnt > i - 1) { left = left - back; back = 2.0 * back; }
            else break;
        }
        back = WERR_(ii);
        for (;;) {
            negcnt = oracle_dlaneg(n, d, lld, right, pivmin, r);
            if (negcnt < i) { right = right + back; back = 2.0 * back; }
            else break;
        }
        width = 0.5 * fabs(left - right);
        tmp = fmax(fabs(left), fabs(right));
        cvrgd = fmax(rtol1 * gap, rtol2 * tmp);
        if (width <= cvrgd || width <= mnwdth) {
            IWORK_(k - 1) = -1;
            if (i == i1 && i < ilast) i1 = i + 1;
            if (prev >= i1 && i <= ilast) IWORK_(2 * prev - 1) = i + 1;
        } else {
            prev = i;
            nint++;
            IWORK_(k - 1) = i + 1;
            IWORK_(k) = negcnt;
        }
        WORK_(k - 1) = left;
        WORK_(k) = right;
    }

    int iter = 0;
    do {
        prev = i1 - 1;
        int i = i1;
        int olnint = nint;
        for (int ip = 1; ip <= olnint; ip++) {
            int k = 2 * i, ii = i - offset;
            rgap = WGAP_(ii);
            lgap = rgap;
            if (ii > 1) lgap = WGAP_(ii - 1);
            gap = fmin(lgap, rgap);
            int next = IWORK_(k - 1);
            left = WORK_(k - 1);
            right = WORK_(k);
            double mid = 0.5 * (left + right);
            width = right - mid;
            tmp = fmax(fabs(left), fabs(right));
            cvrgd = fmax(rtol1 * gap, rtol2 * tmp);
            if (width <= cvrgd || width <= mnwdth || iter == maxitr) {
                nint--;
                IWORK_(k - 1) = 0;
                if (i1 == i) i1 = next;
                else if (prev >= i1) IWORK_(2 * prev - 1) = next;
                i = next;
                continue;
            }
            prev = i;
            int negcnt = oracle_dlaneg(n, d, lld, mid, pivmin, r);
            if (negcnt <= i - 1) WORK_(k - 1) = mid;
            else WORK_(k) = mid;
            i = next;
        }
        iter++;
    } while (nint > 0 && iter <= maxitr);

    for (int i = ifirst; i <= ilast; i++) {
        int k = 2 * i, ii = i - offset;
        if (IWORK_(k - 1) == 0) {
            W_(ii) = 0.5 * (WORK_(k - 1) + WORK_(k));
            WERR_(ii) = WORK_(k) - W_(ii);
        }
    }
    for (int i = ifirst + 1; i <= ilast; i++) {
        int ii = i - offset;
        WGAP_(ii - 1) = fmax(0.0, W_(ii) - WERR_(ii) - W_(ii - 1) - WERR_(ii - 1));
    }
    free(work);
    free(iwork);
#undef W_
#undef WGAP_
#undef WERR_
#undef WORK_
#undef IWORK_
}

/* ------------------------------------------------------------------------------
 * dlar1v: twisted factorisation solve for one eigenvector, odr1v.f:156-369.
 * z points at Z(1) of the block (1-based Z(I) = z[I-1]); work holds 4n doubles.
 * ---------------------------------------------------------------------------- */
void oracle_dlar1v(int n, int b1, int bn, double lambda, const double *d,
                   const double *l, const double *ld, const double *lld,
                   double pivmin, double gaptol, double *z, int wantnc,
                   int *negcnt, double *ztz, double *mingma, int *r,
                   int *isuppz, double *nrminv, double *resid, double *rqcorr,
                   double *work)
{
#define D_(i)    d[(i) - 1]
#define L_(i)    l[(i) - 1]
#define LD_(i)   ld[(i) - 1]
#define LLD_(i)  lld[(i) - 1]
#define Z_(i)    z[(i) - 1]
#define WK(i)    work[(i) - 1]
    int r1, r2, neg1, neg2, i;
    double s, dplus, dminus, tmp;
    int sawnan1, sawnan2;
    if (*r == 0) { r1 = b1; r2 = bn; }
    else { r1 = *r; r2 = *r; }
    const int indlpl = 0, indumn = n, inds = 2 * n + 1, indp = 3 * n + 1;
    if (b1 == 1) WK(inds) = 0.0;
    else WK(inds + b1 - 1) = LLD_(b1 - 1);

    sawnan1 = 0;
    neg1 = 0;
    s = WK(inds + b1 - 1) - lambda;
    for (i = b1; i <= r1 - 1; i++) {
        dplus = D_(i) + s;
        WK(indlpl + i) = LD_(i) / dplus;
        if (dplus < 0.0) neg1++;
        WK(inds + i) = s * WK(indlpl + i) * L_(i);
        s = WK(inds + i) - lambda;
    }
    sawnan1 = isnan(s);
    if (!sawnan1) {
        for (i = r1; i <= r2 - 1; i++) {
            dplus = D_(i) + s;
            WK(indlpl + i) = LD_(i) / dplus;
            WK(inds + i) = s * WK(indlpl + i) * L_(i);
            s = WK(inds + i) - lambda;
        }
        sawnan1 = isnan(s);
    }
    if (sawnan1) {
        neg1 = 0;
        s = WK(inds + b1 - 1) - lambda;
        for (i = b1; i <= r1 - 1; i++) {
            dplus = D_(i) + s;
            if (fabs(dplus) < pivmin) dplus = -pivmin;
            WK(indlpl + i) = LD_(i) / dplus;
            if (dplus < 0.0) neg1++;
            WK(inds + i) = s * WK(indlpl + i) * L_(i);
            if (WK(indlpl + i) == 0.0) WK(inds + i) = LLD_(i);
            s = WK(inds + i) - lambda;
        }
        for (i = r1; i <= r2 - 1; i++) {
            dplus = D_(i) + s;
            if (fabs(dplus) < pivmin) dplus = -pivmin;
            WK(indlpl + i) = LD_(i) / dplus;
            WK(inds + i) = s * WK(indlpl + i) * L_(i);
            if (WK(indlpl + i) == 0.0) WK(inds + i) = LLD_(i);
            s = WK(inds + i) - lambda;
        }
    }

    sawnan2 = 0;
    neg2 = 0;
    WK(indp + bn - 1) = D_(bn) - lambda;
    for (i = bn - 1; i >= r1; i--) {
        dminus = LLD_(i) + WK(indp + i);
        tmp = D_(i) / dminus;
        if (dminus < 0.0) neg2++;
        WK(indumn + i) = L_(i) * tmp;
        WK(indp + i - 1) = WK(indp + i) * tmp - lambda;
    }
    tmp = WK(indp + r1 - 1);
    sawnan2 = isnan(tmp);
    if (sawnan2) {
        neg2 = 0;
        for (i = bn - 1; i >= r1; i--) {
            dminus = LLD_(i) + WK(indp + i);
            if (fabs(dminus) < pivmin) dminus = -pivmin;
            tmp = D_(i) / dminus;
            if (dminus < 0.0) neg2++;
            WK(indumn + i) = L_(i) * tmp;
            WK(indp + i - 1) = WK(indp + i) * tmp - lambda;
            if (tmp == 0.0) WK(indp + i - 1) = D_(i) - lambda;
        }
    }

    *mingma = WK(inds + r1 - 1) + WK(indp + r1 - 1);
    if (*mingma < 0.0) neg1++;
    if (wantnc) *negcnt = neg1 + neg2;
    else *negcnt = -1;
    if (fabs(*mingma) == 0.0) *mingma = EPS * WK(inds + r1 - 1);
    *r = r1;
    for (i = r1; i <= r2 - 1; i++) {
        tmp = WK(inds + i) + WK(indp + i);
        if (tmp == 0.0) tmp = EPS * WK(inds + i);
        if (fabs(tmp) <= fabs(*mingma)) {
            *mingma = tmp;
            *r = i + 1;
        }
    }

    isuppz[0] = b1;
    isuppz[1] = bn;
    Z_(*r) = 1.0;
    *ztz = 1.0;
    int rr = *r;
    if (!sawnan1 && !sawnan2) {
        for (i = rr - 1; i >= b1; i--) {
            Z_(i) = -(WK(indlpl + i) * Z_(i + 1));
            if ((fabs(Z_(i)) + fabs(Z_(i + 1))) * fabs(LD_(i)) < gaptol) {
                Z_(i) = 0.0;
                isuppz[0] = i + 1;
                break;
            }
            *ztz = *ztz + Z_(i) * Z_(i);
        }
    } else {
        for (i = rr - 1; i >= b1; i--) {
            if (Z_(i + 1) == 0.0) Z_(i) = -(LD_(i + 1) / LD_(i)) * Z_(i + 2);
            else Z_(i) = -(WK(indlpl + i) * Z_(i + 1));
            if ((fabs(Z_(i)) + fabs(Z_(i + 1))) * fabs(LD_(i)) < gaptol) {
                Z_(i) = 0.0;
                isuppz[0] = i + 1;
                break;
            }
            *ztz = *ztz + Z_(i) * Z_(i);
        }
    }
    if (!sawnan1 && !sawnan2) {
        for (i = rr; i <= bn - 1; i++) {
            Z_(i + 1) = -(WK(indumn + i) * Z_(i));
            if ((fabs(Z_(i)) + fabs(Z_(i + 1))) * fabs(LD_(i)) < gaptol) {
                Z_(i + 1) = 0.0;
                isuppz[1] = i;
                break;
            }
            *ztz = *ztz + Z_(i + 1) * Z_(i + 1);
        }
    } else {
        for (i = rr; i <= bn - 1; i++) {
            if (Z_(i) == 0.0) Z_(i + 1) = -(LD_(i - 1) / LD_(i)) * Z_(i - 1);
            else Z_(i + 1) = -(WK(indumn + i) * Z_(i));
            if ((fabs(Z_(i)) + fabs(Z_(i + 1))) * fabs(LD_(i)) < gaptol) {
                Z_(i + 1) = 0.0;
                isuppz[1] = i;
                break;
            }
            *ztz = *ztz + Z_(i + 1) * Z_(i + 1);
        }
    }
    tmp = 1.0 / *ztz;
    *nrminv = sqrt(tmp);
    *resid = fabs(*mingma) * *nrminv;
    *rqcorr = *mingma * tmp;
#undef D_
#undef L_
#undef LD_
#undef LLD_
#undef Z_
#undef WK
}

/* ------------------------------------------------------------------------------
 * dlarrf: new RRR L+ D+ L+^T = L D L^T - sigma I for a cluster, odrrf.f:140-384.
 * w, wgap, werr are indexed 1-based from the array start; work holds 2n doubles.
 * Returns INFO (0 = success).
 * ---------------------------------------------------------------------------- */
int oracle_dlarrf(int n, const double *d, const double *l, const double *ld,
                  int clstrt, int clend, const double *w, const double *wgap,
                  const double *werr, double spdiam, double clgapl,
                  double clgapr, double pivmin, double *sigma,
                  double *dplus, double *lplus, double *work)
{
#define W_(i)    w[(i) - 1]
#define WGAP_(i) wgap[(i) - 1]
#define WERR_(i) werr[(i) - 1]
#define DP(i)    dplus[(i) - 1]
#define LP(i)    lplus[(i) - 1]
#define WK(i)    work[(i) - 1]
    const double MAXGROWTH1 = 8.0, MAXGROWTH2 = 8.0;
    const int KTRYMAX = 1, SLEFT = 1, SRIGHT = 2;
    double fact = (double)(1 << KTRYMAX);
    int shift = 0, forcer = 0, nofail = 1, indx = 0;
    int sawnan1, sawnan2, dorrr1, i;
    double clwdth = fabs(W_(clend) - W_(clstrt)) + WERR_(clend) + WERR_(clstrt);
    double avgap = clwdth / (double)(clend - clstrt);
    double mingap = fmin(clgapl, clgapr);
    double lsigma = fmin(W_(clstrt), W_(clend)) - WERR_(clstrt);
    double rsigma = fmax(W_(clstrt), W_(clend)) + WERR_(clend);
    lsigma = lsigma - fabs(lsigma) * 4.0 * EPS;
    rsigma = rsigma + fabs(rsigma) * 4.0 * EPS;
    double ldmax = 0.25 * mingap + 2.0 * pivmin;
    double rdmax = 0.25 * mingap + 2.0 * pivmin;
    double ldelta = fmax(avgap, WGAP_(clstrt)) / fact;
    double rdelta = fmax(avgap, WGAP_(clend - 1)) / fact;
    double s = SAFMIN;
    double smlgrowth = 1.0 / s;
    double fail = (double)(n - 1) * mingap / (spdiam * EPS);
    double fail2 = (double)(n - 1) * mingap / (spdiam * sqrt(EPS));
    double bestshift = lsigma;
    double max1 = 0, max2 = 0, tmp, znm2, prod, oldp, rrr1, rrr2;
    int ktry = 0;
    double growthbound = MAXGROWTH1 * spdiam;

    for (;;) { /* label 5 */
        sawnan1 = 0;
        sawnan2 = 0;
        ldelta = fmin(ldmax, ldelta);
        rdelta = fmin(rdmax, rdelta);
        s = -lsigma;
        DP(1) = d[0] + s;
        if (fabs(DP(1)) < pivmin) { DP(1) = -pivmin; sawnan1 = 1; }
        max1 = fabs(DP(1));
        for (i = 1; i <= n - 1; i++) {
            LP(i) = ld[i - 1] / DP(i);
            s = s * LP(i) * l[i - 1] - lsigma;
            DP(i + 1) = d[i] + s;
            if (fabs(DP(i + 1)) < pivmin) { DP(i + 1) = -pivmin; sawnan1 = 1; }
            max1 = fmax(max1, fabs(DP(i + 1)));
        }
        sawnan1 = sawnan1 || isnan(max1);
        if (forcer || (max1 <= growthbound && !sawnan1)) {
            *sigma = lsigma; shift = SLEFT; goto done;
        }
        s = -rsigma;
        WK(1) = d[0] + s;
        if (fabs(WK(1)) < pivmin) { WK(1) = -pivmin; sawnan2 = 1; }
        max2 = fabs(WK(1));
        for (i = 1; i <= n - 1; i++) {
            WK(n + i) = ld[i - 1] / WK(i);
            s = s * WK(n + i) * l[i - 1] - rsigma;
            WK(i + 1) = d[i] + s;
            if (fabs(WK(i + 1)) < pivmin) { WK(i + 1) = -pivmin; sawnan2 = 1; }
            max2 = fmax(max2, fabs(WK(i + 1)));
        }
        sawnan2 = sawnan2 || isnan(max2);
        if (forcer || (max2 <= growthbound && !sawnan2)) {
            *sigma = rsigma; shift = SRIGHT; goto done;
        }
        if (sawnan1 && sawnan2) goto label50;
        if (!sawnan1) {
            indx = 1;
            if (max1 <= smlgrowth) { smlgrowth = max1; bestshift = lsigma; }
        }
        if (!sawnan2) {
            if (sawnan1 || max2 <= max1) indx = 2;
            if (max2 <= smlgrowth) { smlgrowth = max2; bestshift = rsigma; }
        }
        dorrr1 = (clwdth < mingap / 128.0) && (fmin(max1, max2) < fail2) &&
                 !sawnan1 && !sawnan2;
        if (dorrr1) {
            if (indx == 1) {
                tmp = fabs(DP(n));
                znm2 = 1.0; prod = 1.0; oldp = 1.0;
                for (i = n - 1; i >= 1; i--) {
                    if (prod <= EPS)
                        prod = ((DP(i + 1) * WK(n + i + 1)) / (DP(i) * WK(n + i))) * oldp;
                    else
                        prod = prod * fabs(WK(n + i));
                    oldp = prod;
                    znm2 = znm2 + prod * prod;
                    tmp = fmax(tmp, fabs(DP(i) * prod));
                }
                rrr1 = tmp / (spdiam * sqrt(znm2));
                if (rrr1 <= MAXGROWTH2) { *sigma = lsigma; shift = SLEFT; goto done; }
            } else if (indx == 2) {
                tmp = fabs(WK(n));
                znm2 = 1.0; prod = 1.0; oldp = 1.0;
                for (i = n - 1; i >= 1; i--) {
                    if (prod <= EPS)
                        prod = ((WK(i + 1) * LP(i + 1)) / (WK(i) * LP(i))) * oldp;
                    else
                        prod = prod * fabs(LP(i));
                    oldp = prod;
                    znm2 = znm2 + prod * prod;
                    tmp = fmax(tmp, fabs(WK(i) * prod));
                }
                rrr2 = tmp / (spdiam * sqrt(znm2));
                if (rrr2 <= MAXGROWTH2) { *sigma = rsigma; shift = SRIGHT; goto done; }
            }
        }
    label50:
        if (ktry < KTRYMAX) {
            lsigma = fmax(lsigma - ldelta, lsigma - ldmax);
            rsigma = fmin(rsigma + rdelta, rsigma + rdmax);
            ldelta = 2.0 * ldelta;
            rdelta = 2.0 * rdelta;
            ktry++;
            continue;
        }
        if (smlgrowth < fail || nofail) {
            lsigma = bestshift;
            rsigma = bestshift;
            forcer = 1;
            continue;
        }
        return 1;
    }
done:
    if (shift == SRIGHT) {
        for (i = 1; i <= n; i++) DP(i) = WK(i);
        for (i = 1; i <= n - 1; i++) LP(i) = WK(n + i);
    }
    return 0;
#undef W_
#undef WGAP_
#undef WERR_
#undef DP
#undef LP
#undef WK
}

/* ------------------------------------------------------------------------------
 * dlaebz, scalar path (NBMIN = 0), IJOB = 1, 2, 3: odebz.f:237-536.
 * AB, NAB are MMAX x 2 column-major; C, NVAL have MMAX entries.
 * ---------------------------------------------------------------------------- */
static int o_dlaebz(int ijob, int nitmax, int n, int mmax, int minp,
                    double abstol, double reltol, double pivmin,
                    const double *d, const double *e2, int *nval, double *ab,
                    double *c, int *mout, int *nab)
{
#define AB_(ji, jp)  ab[((jp) - 1) * mmax + (ji) - 1]
#define NAB_(ji, jp) nab[((jp) - 1) * mmax + (ji) - 1]
#define C_(ji)       c[(ji) - 1]
#define NVAL_(ji)    nval[(ji) - 1]
    int info = 0;
    if (ijob == 1) {
        *mout = 0;
        for (int ji = 1; ji <= minp; ji++) {
            for (int jp = 1; jp <= 2; jp++) {
                double tmp1 = d[0] - AB_(ji, jp);
                if (fabs(tmp1) < pivmin) tmp1 = -pivmin;
                NAB_(ji, jp) = 0;
                if (tmp1 <= 0.0) NAB_(ji, jp) = 1;
                for (int j = 2; j <= n; j++) {
                    tmp1 = d[j - 1] - e2[j - 2] / tmp1 - AB_(ji, jp);
                    if (fabs(tmp1) < pivmin) tmp1 = -pivmin;
                    if (tmp1 <= 0.0) NAB_(ji, jp) = NAB_(ji, jp) + 1;
                }
            }
            *mout = *mout + NAB_(ji, 2) - NAB_(ji, 1);
        }
        return 0;
    }
    int kf = 1, kl = minp;
    if (ijob == 2)
        for (int ji = 1; ji <= minp; ji++) C_(ji) = 0.5 * (AB_(ji, 1) + AB_(ji, 2));
    for (int jit = 1; jit <= nitmax; jit++) {
        int klnew = kl;
        for (int ji = kf; ji <= kl; ji++) {
            double tmp1 = C_(ji);
            double tmp2 = d[0] - tmp1;
            int itmp1 = 0;
            if (tmp2 <= pivmin) { itmp1 = 1; tmp2 = fmin(tmp2, -pivmin); }
            for (int j = 2; j <= n; j++) {
                tmp2 = d[j - 1] - e2[j - 2] / tmp2 - tmp1;
                if (tmp2 <= pivmin) { itmp1++; tmp2 = fmin(tmp2, -pivmin); }
            }
            if (ijob <= 2) {
                itmp1 = imin(NAB_(ji, 2), imax(NAB_(ji, 1), itmp1));
                if (itmp1 == NAB_(ji, 2)) {
                    AB_(ji, 2) = tmp1;
                } else if (itmp1 == NAB_(ji, 1)) {
                    AB_(ji, 1) = tmp1;
                } else if (klnew < mmax) {
                    klnew++;
                    AB_(klnew, 2) = AB_(ji, 2);
                    NAB_(klnew, 2) = NAB_(ji, 2);
                    AB_(klnew, 1) = tmp1;
                    NAB_(klnew, 1) = itmp1;
                    AB_(ji, 2) = tmp1;
                    NAB_(ji, 2) = itmp1;
                } else {
                    return mmax + 1;
                }
            } else {
                if (itmp1 <= NVAL_(ji)) { AB_(ji, 1) = tmp1; NAB_(ji, 1) = itmp1; }
                if (itmp1 >= NVAL_(ji)) { AB_(ji, 2) = tmp1; NAB_(ji, 2) = itmp1; }
            }
        }
        kl = klnew;
        int kfnew = kf;
        for (int ji = kf; ji <= kl; ji++) {
            double tmp1 = fabs(AB_(ji, 2) - AB_(ji, 1));
            double tmp2 = fmax(fabs(AB_(ji, 2)), fabs(AB_(ji, 1)));
            if (tmp1 < fmax(fmax(abstol, pivmin), reltol * tmp2) ||
                NAB_(ji, 1) >= NAB_(ji, 2)) {
                if (ji > kfnew) {
                    double t1 = AB_(ji, 1), t2 = AB_(ji, 2);
                    int i1 = NAB_(ji, 1), i2 = NAB_(ji, 2);
                    AB_(ji, 1) = AB_(kfnew, 1);
                    AB_(ji, 2) = AB_(kfnew, 2);
                    NAB_(ji, 1) = NAB_(kfnew, 1);
                    NAB_(ji, 2) = NAB_(kfnew, 2);
                    AB_(kfnew, 1) = t1;
                    AB_(kfnew, 2) = t2;
                    NAB_(kfnew, 1) = i1;
                    NAB_(kfnew, 2) = i2;
                    if (ijob == 3) {
                        int it = NVAL_(ji);
                        NVAL_(ji) = NVAL_(kfnew);
                        NVAL_(kfnew) = it;
                    }
                }
                kfnew++;
            }
        }
        kf = kfnew;
        for (int ji = kf; ji <= kl; ji++) C_(ji) = 0.5 * (AB_(ji, 1) + AB_(ji, 2));
        if (kf > kl) break;
    }
    info = imax(kl + 1 - kf, 0);
    *mout = kl;
    return info;
#undef AB_
#undef NAB_
#undef C_
#undef NVAL_
}

/* ------------------------------------------------------------------------------
 * dlarrd with ORDER = 'B': eigenvalue approximations by bisection on T,
 * odrrd.f:241-722 (range I bracket :343-393, block loop :415-574, discard
 * :576-690).  gers holds 2n Gershgorin bounds.  Returns INFO.
 * ---------------------------------------------------------------------------- */
int o_dlarrd(char range, int n, double vl, double vu, int il, int iu,
                    const double *gers, double reltol, const double *d,
                    const double *e2, double pivmin, int nsplit,
                    const int *isplit, int *m, double *w, double *werr,
                    double *wl_out, double *wu_out, int *iblock, int *indexw)
{
    const double FUDGE = 2.0;
    int irange = (range == 'A') ? 1 : (range == 'V') ? 2 : 3;
    int info = 0, ncnvrg = 0, toofew = 0;
    *m = 0;
    if (n == 0) return 0;
    if (irange == 3 && il == 1 && iu == n) irange = 1;
    if (n == 1) {
        if (irange == 1 || (irange == 2 && d[0] > vl && d[0] <= vu) ||
            (irange == 3 && il == 1 && iu == 1)) {
            *m = 1; w[0] = d[0]; werr[0] = 0.0; iblock[0] = 1; indexw[0] = 1;
        }
        return 0;
    }
    double gl = d[0], gu = d[0];
    for (int i = 1; i <= n; i++) {
        gl = fmin(gl, gers[2 * i - 2]);
        gu = fmax(gu, gers[2 * i - 1]);
    }
    double tnorm = fmax(fabs(gl), fabs(gu));
    gl = gl - FUDGE * tnorm * EPS * n - FUDGE * 2.0 * pivmin;
    gu = gu + FUDGE * tnorm * EPS * n + FUDGE * 2.0 * pivmin;
    double rtoli = reltol;
    double atoli = FUDGE * 2.0 * SAFMIN + FUDGE * 2.0 * pivmin;
    double wl = 0, wu = 0, wlu = 0, wul = 0;
    int nwl = 0, nwu = 0;

    double *work = (double *)malloc(sizeof(double) * (size_t)(3 * n + 8));
    int *iwork = (int *)malloc(sizeof(int) * (size_t)(2 * n + 8));
    if (irange == 3) {
        int itmax = (int)((log(tnorm + pivmin) - log(pivmin)) / log(2.0)) + 2;
        double ab[4] = {gl, gl, gu, gu}, c[2] = {gl, gu};
        int nab[4] = {-1, -1, n + 1, n + 1}, nval[2] = {il - 1, iu}, iout;
        int iinfo = o_dlaebz(3, itmax, n, 2, 2, atoli, rtoli, pivmin, d, e2, nval,
                             ab, c, &iout, nab);
        if (iinfo != 0) { free(work); free(iwork); return iinfo; }
        if (nval[1] == iu) {
            wl = ab[0]; wlu = ab[2]; nwl = nab[0];
            wu = ab[3]; wul = ab[1]; nwu = nab[3];
        } else {
            wl = ab[1]; wlu = ab[3]; nwl = nab[1];
            wu = ab[2]; wul = ab[0]; nwu = nab[2];
        }
        if (nwl < 0 || nwl >= n || nwu < 1 || nwu > n) {
            free(work); free(iwork); return 4;
        }
    } else if (irange == 2) {
        wl = vl; wu = vu;
    } else {
        wl = gl; wu = gu;
    }

    int mm = 0, iend = 0;
    nwl = (irange == 3) ? nwl : 0;
    nwu = (irange == 3) ? nwu : 0;
    /* odrrd.f:413-415 resets NWL and NWU before the block loop */
    int nwl_b = 0, nwu_b = 0;
    for (int jblk = 1; jblk <= nsplit; jblk++) {
        int ioff = iend;
        int ibegin = ioff + 1;
        iend = isplit[jblk - 1];
        int in = iend - ioff;
        if (in == 1) {
            if (wl >= d[ibegin - 1] - pivmin) nwl_b++;
            if (wu >= d[ibegin - 1] - pivmin) nwu_b++;
            if (irange == 1 || (wl < d[ibegin - 1] - pivmin && wu >= d[ibegin - 1] - pivmin)) {
                mm++;
                w[mm - 1] = d[ibegin - 1];
                werr[mm - 1] = 0.0;
                iblock[mm - 1] = jblk;
                indexw[mm - 1] = 1;
            }
        } else {
            double bgu = d[ibegin - 1], bgl = d[ibegin - 1];
            for (int j = ibegin; j <= iend; j++) {
                bgl = fmin(bgl, gers[2 * j - 2]);
                bgu = fmax(bgu, gers[2 * j - 1]);
            }
            bgl = bgl - FUDGE * tnorm * EPS * in - FUDGE * pivmin;
            bgu = bgu + FUDGE * tnorm * EPS * in + FUDGE * pivmin;
            if (irange > 1) {
                if (bgu < wl) { nwl_b += in; nwu_b += in; continue; }
                bgl = fmax(bgl, wl);
                bgu = fmin(bgu, wu);
                if (bgl >= bgu) continue;
            }
            /* AB(IN,2) in work[0..2in), C in work[2in..3in), NAB(IN,2) in iwork */
            double *ab = work, *c = work + 2 * in;
            int *nab = iwork;
            ab[0] = bgl;
            ab[in] = bgu;
            int im, iout, dummy_nval = 0;
            o_dlaebz(1, 0, in, in, 1, atoli, rtoli, pivmin, d + ibegin - 1,
                     e2 + ibegin - 1, &dummy_nval, ab, c, &im, nab);
            nwl_b += nab[0];
            nwu_b += nab[in];
            int iwoff = mm - nab[0];
            int itmax = (int)((log(bgu - bgl + pivmin) - log(pivmin)) / log(2.0)) + 2;
            int iinfo = o_dlaebz(2, itmax, in, in, 1, atoli, rtoli, pivmin,
                                 d + ibegin - 1, e2 + ibegin - 1, &dummy_nval,
                                 ab, c, &iout, nab);
            if (iinfo < 0 || iinfo > in) { free(work); free(iwork); return iinfo; }
            for (int j = 1; j <= iout; j++) {
                double tmp1 = 0.5 * (ab[j - 1] + ab[j - 1 + in]);
                double tmp2 = 0.5 * fabs(ab[j - 1] - ab[j - 1 + in]);
                int ib;
                if (j > iout - iinfo) { ncnvrg = 1; ib = -jblk; }
                else ib = jblk;
                for (int je = nab[j - 1] + 1 + iwoff; je <= nab[j - 1 + in] + iwoff; je++) {
                    w[je - 1] = tmp1;
                    werr[je - 1] = tmp2;
                    indexw[je - 1] = je - iwoff;
                    iblock[je - 1] = ib;
                }
            }
            mm += im;
        }
    }

    if (irange == 3) {
        int idiscl = il - 1 - nwl_b;
        int idiscu = nwu_b - iu;
        if (idiscl > 0) {
            int im = 0;
            for (int je = 1; je <= mm; je++) {
                if (w[je - 1] <= wlu && idiscl > 0) idiscl--;
                else {
                    im++;
                    w[im - 1] = w[je - 1]; werr[im - 1] = werr[je - 1];
                    indexw[im - 1] = indexw[je - 1]; iblock[im - 1] = iblock[je - 1];
                }
            }
            mm = im;
        }
        if (idiscu > 0) {
            int im = mm + 1;
            for (int je = mm; je >= 1; je--) {
                if (w[je - 1] >= wul && idiscu > 0) idiscu--;
                else {
                    im--;
                    w[im - 1] = w[je - 1]; werr[im - 1] = werr[je - 1];
                    indexw[im - 1] = indexw[je - 1]; iblock[im - 1] = iblock[je - 1];
                }
            }
            int jee = 0;
            for (int je = im; je <= mm; je++) {
                jee++;
                w[jee - 1] = w[je - 1]; werr[jee - 1] = werr[je - 1];
                indexw[jee - 1] = indexw[je - 1]; iblock[jee - 1] = iblock[je - 1];
            }
            mm = mm - im + 1;
        }
        if (idiscl > 0 || idiscu > 0) {
            if (idiscl > 0) {
                double wkill = wu;
                for (int jd = 1; jd <= idiscl; jd++) {
                    int iw = 0;
                    for (int je = 1; je <= mm; je++)
                        if (iblock[je - 1] != 0 && (w[je - 1] < wkill || iw == 0)) {
                            iw = je; wkill = w[je - 1];
                        }
                    iblock[iw - 1] = 0;
                }
            }
            if (idiscu > 0) {
                double wkill = wl;
                for (int jd = 1; jd <= idiscu; jd++) {
                    int iw = 0;
                    for (int je = 1; je <= mm; je++)
                        if (iblock[je - 1] != 0 && (w[je - 1] >= wkill || iw == 0)) {
                            iw = je; wkill = w[je - 1];
                        }
                    iblock[iw - 1] = 0;
                }
            }
            int im = 0;
            for (int je = 1; je <= mm; je++)
                if (iblock[je - 1] != 0) {
                    im++;
                    w[im - 1] = w[je - 1]; werr[im - 1] = werr[je - 1];
                    indexw[im - 1] = indexw[je - 1]; iblock[im - 1] = iblock[je - 1];
                }
            mm = im;
        }
        if (idiscl < 0 || idiscu < 0) toofew = 1;
    }
    if ((irange == 1 && mm != n) || (irange == 3 && mm != iu - il + 1)) toofew = 1;
    *m = mm;
    *wl_out = wl;
    *wu_out = wu;
    free(work);
    free(iwork);
    info = 0;
    if (ncnvrg) info += 1;
    if (toofew) info += 2;
    return info;
}

/* dlarrc with JOBT = 'T': counts on T at vl and vu, odrrc.f:95-113 */
void o_dlarrc_T(int n, double vl, double vu, const double *d,
                       const double *e, int *lcnt, int *rcnt)
{
    double lpivot = d[0] - vl, rpivot = d[0] - vu;
    *lcnt = 0; *rcnt = 0;
    if (lpivot <= 0.0) (*lcnt)++;
    if (rpivot <= 0.0) (*rcnt)++;
    for (int i = 1; i <= n - 1; i++) {
        double tmp = e[i - 1] * e[i - 1];
        lpivot = (d[i] - vl) - tmp / lpivot;
        rpivot = (d[i] - vu) - tmp / rpivot;
        if (lpivot <= 0.0) (*lcnt)++;
        if (rpivot <= 0.0) (*rcnt)++;
    }
}

/* dlarra: splitting points, odrra.f:90-123 (isplit 1-based ends) */
void o_dlarra(int n, const double *d, double *e, double *e2, double spltol,
                     double tnrm, int *nsplit, int *isplit)
{
    *nsplit = 1;
    if (spltol < 0.0) {
        double tmp1 = fabs(spltol) * tnrm;
        for (int i = 1; i <= n - 1; i++)
            if (fabs(e[i - 1]) <= tmp1) {
                e[i - 1] = 0.0; e2[i - 1] = 0.0;
                isplit[*nsplit - 1] = i; (*nsplit)++;
            }
    } else {
        for (int i = 1; i <= n - 1; i++)
            if (fabs(e[i - 1]) <= spltol * sqrt(fabs(d[i - 1])) * sqrt(fabs(d[i]))) {
                e[i - 1] = 0.0; e2[i - 1] = 0.0;
                isplit[*nsplit - 1] = i; (*nsplit)++;
            }
    }
    isplit[*nsplit - 1] = n;
}

/* dlarrr: test for relative accuracy, odrrr.f:60-121 (0 = yes) */
int o_dlarrr(int n, const double *d, const double *e)
{
    const double RELCOND = 0.999;
    double smlnum = SAFMIN / EPS, rmin = sqrt(smlnum);
    int yesrel = 1;
    double offdig = 0.0, tmp = sqrt(fabs(d[0]));
    if (tmp < rmin) yesrel = 0;
    for (int i = 2; i <= n && yesrel; i++) {
        double tmp2 = sqrt(fabs(d[i - 1]));
        if (tmp2 < rmin) { yesrel = 0; break; }
        double offdig2 = fabs(e[i - 2]) / (tmp * tmp2);
        if (offdig + offdig2 >= RELCOND) { yesrel = 0; break; }
        tmp = tmp2;
        offdig = offdig2;
    }
    return yesrel ? 0 : 1;
}

/* dlarrj: bisection on T to relative accuracy RTOL, odrrj.f:100-281 */
void o_dlarrj(int n, const double *d, const double *e2, int ifirst,
                     int ilast, double rtol, int offset, double *w,
                     double *werr, double pivmin, double spdiam)
{
    double *work = (double *)malloc(sizeof(double) * (size_t)(2 * ilast + 2));
    int *iwork = (int *)malloc(sizeof(int) * (size_t)(2 * ilast + 2));
    int maxitr = (int)((log(spdiam + pivmin) - log(pivmin)) / log(2.0)) + 2;
    int i1 = ifirst, i2 = ilast, nint = 0, prev = 0;
    for (int i = i1; i <= i2; i++) {
        int k = 2 * i, ii = i - offset;
        double left = w[ii - 1] - werr[ii - 1];
        double mid = w[ii - 1];
        double right = w[ii - 1] + werr[ii - 1];
        double width = right - mid;
        double tmp = fmax(fabs(left), fabs(right));
        if (width < rtol * tmp) {
            iwork[k - 1] = -1;
            if (i == i1 && i < i2) i1 = i + 1;
            if (prev >= i1 && i <= i2) iwork[2 * prev - 1] = i + 1;
        } else {
            prev = i;
            double fac = 1.0;
            int cnt;
            for (;;) {
                cnt = 0;
                double s = left, dplus = d[0] - s;
                if (dplus < 0.0) cnt++;
                for (int j = 2; j <= n; j++) {
                    dplus = d[j - 1] - s - e2[j - 2] / dplus;
                    if (dplus < 0.0) cnt++;
                }
                if (cnt > i - 1) { left = left - werr[ii - 1] * fac; fac = 2.0 * fac; }
                else break;
            }
            fac = 1.0;
            for (;;) {
                cnt = 0;
                double s = right, dplus = d[0] - s;
                if (dplus < 0.0) cnt++;
                for (int j = 2; j <= n; j++) {
                    dplus = d[j - 1] - s - e2[j - 2] / dplus;
                    if (dplus < 0.0) cnt++;
                }
                if (cnt < i) { right = right + werr[ii - 1] * fac; fac = 2.0 * fac; }
                else break;
            }
            nint++;
            iwork[k - 1] = i + 1;
            iwork[k] = cnt;
        }
        work[k - 1] = left;
        work[k] = right;
    }
    int savi1 = i1, iter = 0;
    do {
        prev = i1 - 1;
        int i = i1, olnint = nint;
        for (int p = 1; p <= olnint; p++) {
            int k = 2 * i;
            int next = iwork[k - 1];
            double left = work[k - 1], right = work[k];
            double mid = 0.5 * (left + right);
            double width = right - mid;
            double tmp = fmax(fabs(left), fabs(right));
            if (width < rtol * tmp || iter == maxitr) {
                nint--;
                iwork[k - 1] = 0;
                if (i1 == i) i1 = next;
                else if (prev >= i1) iwork[2 * prev - 1] = next;
                i = next;
                continue;
            }
            prev = i;
            int cnt = 0;
            double s = mid, dplus = d[0] - s;
            if (dplus < 0.0) cnt++;
            for (int j = 2; j <= n; j++) {
                dplus = d[j - 1] - s - e2[j - 2] / dplus;
                if (dplus < 0.0) cnt++;
            }
            if (cnt <= i - 1) work[k - 1] = mid;
            else work[k] = mid;
            i = next;
        }
        iter++;
    } while (nint > 0 && iter <= maxitr);
    for (int i = savi1; i <= ilast; i++) {
        int k = 2 * i, ii = i - offset;
        if (iwork[k - 1] == 0) {
            w[ii - 1] = 0.5 * (work[k - 1] + work[k]);
            werr[ii - 1] = work[k] - w[ii - 1];
        }
    }
    free(work);
    free(iwork);
}

/* dlaev2 (odev2.f:70-170): eigen-decomposition of [[a,b],[b,c]] */
void o_dlaev2(double a, double b, double c, double *rt1, double *rt2,
                     double *cs1, double *sn1)
{
    double sm = a + c, df = a - c, adf = fabs(df), tb = b + b, ab = fabs(tb);
    double acmx, acmn, rt, cs, ct, tn, acs;
    int sgn1, sgn2;
    if (fabs(a) > fabs(c)) { acmx = a; acmn = c; }
    else { acmx = c; acmn = a; }
    if (adf > ab) rt = adf * sqrt(1.0 + (ab / adf) * (ab / adf));
    else if (adf < ab) rt = ab * sqrt(1.0 + (adf / ab) * (adf / ab));
    else rt = ab * sqrt(2.0);
    if (sm < 0.0) { *rt1 = 0.5 * (sm - rt); sgn1 = -1; *rt2 = (acmx / *rt1) * acmn - (b / *rt1) * b; }
    else if (sm > 0.0) { *rt1 = 0.5 * (sm + rt); sgn1 = 1; *rt2 = (acmx / *rt1) * acmn - (b / *rt1) * b; }
    else { *rt1 = 0.5 * rt; *rt2 = -0.5 * rt; sgn1 = 1; }
    if (df >= 0.0) { cs = df + rt; sgn2 = 1; }
    else { cs = df - rt; sgn2 = -1; }
    acs = fabs(cs);
    if (acs > ab) {
        ct = -tb / cs;
        *sn1 = 1.0 / sqrt(1.0 + ct * ct);
        *cs1 = ct * *sn1;
    } else {
        if (ab == 0.0) { *cs1 = 1.0; *sn1 = 0.0; }
        else {
            tn = -cs / tb;
            *cs1 = 1.0 / sqrt(1.0 + tn * tn);
            *sn1 = tn * *cs1;
        }
    }
    if (sgn1 == sgn2) { tn = *cs1; *cs1 = -*sn1; *sn1 = tn; }
}

/* ==============================================================================
 * Eigenvalue stage: mrrr_val (mrrr_val.c:110-639), bisection branch only.
 * ============================================================================ */
typedef struct {
    int n;
    double *D, *E, *E2;
    int nsplit;
    int *isplit;        /* 1-based block ends */
    double spdiam;      /* Dstruct->spdiam (mrrr_val.c:236) */
    double *W, *Werr, *Wgap, *gersch, *Wshifted;
    int *iblock, *Windex, *Zindex;
    int m;
    double pivmin, rtol1, rtol2, bsrtol, RQtol, split;
} o_state;

static int o_compute_root_rrr(o_state *st, int ibgn, int iend, int wbgn,
                              int wend, int ilow, int iupp, double gl,
                              double gu, double vl, double vu, int *sgndefp,
                              double *sigmap, double *spdiamp);

static int o_mrrr_val(char jobz, char range, o_state *st, double *vlp,
                      double *vup, int *ilp, int *iup)
{
    int n = st->n;
    double *D = st->D, *E = st->E, *E2 = st->E2;
    double *W = st->W, *Werr = st->Werr, *Wgap = st->Wgap, *gersch = st->gersch;
    int *iblock = st->iblock, *Windex = st->Windex, *isplit = st->isplit;
    int irange = (range == 'A') ? 1 : (range == 'V') ? 2 : 3;
    if (irange == 3 && *ilp == 1 && *iup == n) irange = 1;

    if (n == 1) { /* mrrr_val.c:193-211 */
        if (irange == 1 || (irange == 2 && D[0] > *vlp && D[0] <= *vup) ||
            (irange == 3 && *ilp == 1 && *iup == 1)) {
            st->m = 1; W[0] = D[0]; Werr[0] = 0.0; Wgap[0] = 0.0;
            iblock[0] = 1; Windex[0] = 1; gersch[0] = D[0]; gersch[1] = D[0];
        } else st->m = 0;
        E[0] = 0.0;
        st->nsplit = 1; isplit[0] = 1;
        return 0;
    }
    for (int i = 0; i < n; i++) E2[i] = E[i] * E[i];
    double gl = D[0], gu = D[0], eold = 0.0, emax = 0.0;
    E[n - 1] = 0.0;
    for (int i = 0; i < n; i++) { /* mrrr_val.c:221-231 */
        double eabs = fabs(E[i]);
        if (eabs >= emax) emax = eabs;
        double tmp1 = eabs + eold;
        gersch[2 * i] = D[i] - tmp1;
        gl = fmin(gl, gersch[2 * i]);
        gersch[2 * i + 1] = D[i] + tmp1;
        gu = fmax(gu, gersch[2 * i + 1]);
        eold = eabs;
    }
    st->pivmin = DBL_MIN * fmax(1.0, emax * emax);
    st->spdiam = gu - gl;
    double tnorm = fmax(fabs(gl), fabs(gu));
    gl = gl - 2 * tnorm * EPS * n - 4 * st->pivmin;
    gu = gu + 2 * tnorm * EPS * n + 4 * st->pivmin;
    if (irange == 2 && *vlp < gl && *vup >= gu) irange = 1;
    o_dlarra(n, D, E, E2, st->split, st->spdiam, &st->nsplit, isplit);

    if (irange == 1) { *ilp = 1; *iup = n; }
    else if (irange == 2) {
        double ab[2] = {*vlp, *vup}, c[1];
        int nab[2], mout, nv = 0;
        o_dlaebz(1, 0, n, 1, 1, 0.0, 0.0, st->pivmin, D, E2, &nv, ab, c, &mout, nab);
        *ilp = nab[0] + 1;
        *iup = nab[1];
    }
    int nvals = *iup - *ilp + 1;
    st->m = nvals;
    if (jobz == 'C') return 0;

    double wl, wu, wlu = 0, wul = 0;
    int nwl = 0, nwu = 0;
    if (irange == 3) { /* mrrr_val.c:293-319 */
        int itmax = (int)((log(tnorm + st->pivmin) - log(st->pivmin)) / log(2.0) + 2);
        double atoli = 2 * (2 * DBL_MIN + 2 * st->pivmin);
        double ab[4] = {gl, gl, gu, gu}, c[2] = {gl, gu};
        int nab[4] = {-1, -1, n + 1, n + 1}, nval[2] = {*ilp - 1, *iup}, mout;
        o_dlaebz(3, itmax, n, 2, 2, atoli, st->RQtol, st->pivmin, D, E2, nval, ab, c, &mout, nab);
        if (nval[1] == *iup) {
            wl = ab[0]; wlu = ab[2]; wu = ab[3]; wul = ab[1]; nwl = nab[0]; nwu = nab[3];
        } else {
            wl = ab[1]; wlu = ab[3]; wu = ab[2]; wul = ab[0]; nwu = nab[2]; nwl = nab[1];
        }
        *vlp = wl; *vup = wu;
    } else if (irange == 2) {
        wl = *vlp; wu = *vup;
    } else {
        wl = gl; wu = gu; *vlp = gl; *vup = gu;
    }

    double *Wcopy = (double *)malloc(sizeof(double) * (size_t)n);
    int mp = 0, ibgn = 0, wbgn = 0;
    for (int jbl = 0; jbl < st->nsplit; jbl++) { /* mrrr_val.c:340-512 */
        int iend = isplit[jbl] - 1;
        int nbl = iend - ibgn + 1;
        if (nbl == 1) {
            if (irange == 1 || (irange == 2 && D[ibgn] > wl && D[ibgn] <= wu) ||
                (irange == 3 && (D[ibgn] + st->pivmin) > wl && (D[ibgn] - st->pivmin) <= wu)) {
                W[mp] = D[ibgn]; Wcopy[mp] = D[ibgn]; Werr[mp] = 0.0;
                iblock[mp] = jbl + 1; Windex[mp] = 1;
                mp++; wbgn++;
            }
            E[iend] = 0.0;
            ibgn = iend + 1;
            continue;
        }
        double bgl = D[ibgn], bgu = D[ibgn];
        for (int i = ibgn; i <= iend; i++) {
            bgl = fmin(bgl, gersch[2 * i]);
            bgu = fmax(bgu, gersch[2 * i + 1]);
        }
        double spdiam = bgu - bgl;
        bgl = bgl - 2 * tnorm * EPS * nbl - 4 * st->pivmin;
        bgu = bgu + 2 * tnorm * EPS * nbl + 4 * st->pivmin;
        int ilow, iupp, wend;
        if (irange == 1) { ilow = 1; iupp = nbl; }
        else {
            double ab[2] = {fmax(wl, bgl), fmin(wu, bgu)}, c[1];
            int nab[2], mout, nv = 0;
            o_dlaebz(1, 0, nbl, 1, 1, 0.0, 0.0, st->pivmin, D + ibgn, E2 + ibgn, &nv, ab, c, &mout, nab);
            ilow = nab[0] + 1;
            iupp = nab[1];
        }
        int mbl = iupp - ilow + 1;
        mp += mbl;
        if (mbl == 0) {
            E[iend] = 0.0;
            ibgn = iend + 1;
            continue;
        }
        wend = wbgn + mbl - 1;
        for (int i = wbgn; i <= wend; i++) iblock[i] = jbl + 1;
        { /* find_eigval_irreducible, sequential case (mrrr_val.c:1279-1300) */
            int isp = nbl, mo;
            double dwl, dwu;
            int *tib = (int *)malloc(sizeof(int) * (size_t)nbl);
            o_dlarrd('I', nbl, 0.0, 0.0, ilow, iupp, gersch + 2 * ibgn, st->bsrtol,
                     D + ibgn, E2 + ibgn, st->pivmin, 1, &isp, &mo, W + wbgn,
                     Werr + wbgn, &dwl, &dwu, tib, Windex + wbgn);
            free(tib);
        }
        for (int i = wbgn; i <= wend; i++) Wcopy[i] = W[i];
        for (int i = wbgn; i < wend; i++)
            Wgap[i] = fmax(0.0, (W[i + 1] - Werr[i + 1]) - (W[i] + Werr[i]));
        Wgap[wend] = fmax(0.0, wu - (W[wend] + Werr[wend]));

        int sgndef;
        double sigma;
        int info = o_compute_root_rrr(st, ibgn, iend, wbgn, wend, ilow, iupp, bgl, bgu,
                                      *vlp, *vup, &sgndef, &sigma, &spdiam);
        if (info != 0) { free(Wcopy); return 2; }
        { /* refine_eigval_approx, sequential case (mrrr_val.c:954-1087) */
            for (int i = wbgn; i <= wend; i++) {
                W[i] -= sigma;
                Werr[i] += fabs(W[i]) * DBL_EPSILON;
            }
            double *DE2 = (double *)malloc(sizeof(double) * (size_t)nbl);
            for (int i = ibgn; i < iend; i++) DE2[i - ibgn] = D[i] * E[i] * E[i];
            DE2[nbl - 1] = 0.0;
            oracle_dlarrb(nbl, D + ibgn, DE2, ilow, iupp, st->rtol1, st->rtol2, ilow - 1,
                          W + wbgn, Wgap + wbgn, Werr + wbgn, st->pivmin, spdiam, nbl);
            free(DE2);
            Wgap[wend] = fmax(0.0, (*vup - sigma) - (W[wend] + Werr[wend]));
        }
        ibgn = iend + 1;
        wbgn = wend + 1;
    }

    if (mp > nvals) { /* mrrr_val.c:520-626 */
        int idiscl = *ilp - 1 - nwl, idiscu = nwu - *iup, im;
        if (idiscl > 0) {
            im = 0;
            for (int je = 0; je < mp; je++) {
                if (Wcopy[je] < wlu && idiscl > 0) idiscl--;
                else {
                    W[im] = W[je]; Wcopy[im] = Wcopy[je]; Werr[im] = Werr[je];
                    Wgap[im] = Wgap[je]; Windex[im] = Windex[je]; iblock[im] = iblock[je];
                    im++;
                }
            }
            mp = im;
        }
        if (idiscu > 0) {
            im = mp;
            for (int je = mp - 1; je >= 0; je--) {
                if (Wcopy[je] > wul && idiscu > 0) idiscu--;
                else {
                    im--;
                    W[im] = W[je]; Wcopy[im] = Wcopy[je]; Werr[im] = Werr[je];
                    Wgap[im] = Wgap[je]; Windex[im] = Windex[je]; iblock[im] = iblock[je];
                }
            }
            int jee = 0;
            for (int je = im; je < mp; je++) {
                W[jee] = W[je]; Wcopy[jee] = Wcopy[je]; Werr[jee] = Werr[je];
                Wgap[jee] = Wgap[je]; Windex[jee] = Windex[je]; iblock[jee] = iblock[je];
                jee++;
            }
            mp -= im;
        }
        if (idiscl > 0 || idiscu > 0) {
            if (idiscl > 0) {
                double wkill = wu;
                for (int jd = 0; jd < idiscl; jd++) {
                    int iw = -1;
                    for (int je = 0; je < mp; je++)
                        if (iblock[je] != 0 && (Wcopy[je] < wkill || iw == -1)) {
                            iw = je; wkill = Wcopy[je];
                        }
                    iblock[iw] = 0;
                }
            }
            if (idiscu > 0) {
                double wkill = wl;
                for (int jd = 0; jd < idiscu; jd++) {
                    int iw = -1;
                    for (int je = 0; je < mp; je++)
                        if (iblock[je] != 0 && (Wcopy[je] > wkill || iw == -1)) {
                            iw = je; wkill = Wcopy[je];
                        }
                    iblock[iw] = 0;
                }
            }
            im = 0;
            for (int je = 0; je < mp; je++)
                if (iblock[je] != 0) {
                    W[im] = W[je]; Wcopy[im] = Wcopy[je]; Werr[im] = Werr[je];
                    Wgap[im] = Wgap[je]; Windex[im] = Windex[je]; iblock[im] = iblock[je];
                    im++;
                }
            mp = im;
        }
    }
    free(Wcopy);
    st->m = mp;
    if (mp < nvals) return 3;
    return 0;
}

/* compute_root_rrr, bisection case (mrrr_val.c:717-948) */
static int o_compute_root_rrr(o_state *st, int ibgn, int iend, int wbgn,
                              int wend, int ilow, int iupp, double gl,
                              double gu, double vl, double vu, int *sgndefp,
                              double *sigmap, double *spdiamp)
{
    int nbl = iend - ibgn + 1, mbl = wend - wbgn + 1;
    double *D = st->D, *E = st->E, *W = st->W, *Werr = st->Werr, *Wgap = st->Wgap;
    double pivmin = st->pivmin;
    double isleft = fmax(gl, (W[wbgn] - Werr[wbgn]) - HUNDRED * EPS * fabs(W[wbgn] - Werr[wbgn]));
    double isright = fmin(gu, (W[wend] + Werr[wend]) + HUNDRED * EPS * fabs(W[wend] - Werr[wend]));
    double spdiam = gu - gl;
    double tmp0 = fmin(isright, vu) - fmax(isleft, vl);
    double s1 = fmax(isleft, vl) + FOURTH * tmp0;
    double s2 = fmin(isright, vu) - FOURTH * tmp0;
    int negcnt_lft = 0, negcnt_rgt = 0, sgndef;
    double sigma, tau;
    if (mbl > 1) o_dlarrc_T(nbl, s1, s2, D + ibgn, E + ibgn, &negcnt_lft, &negcnt_rgt);
    if (mbl == 1) { sigma = gl; sgndef = 1; }
    else if ((negcnt_lft - ilow) >= (iupp - negcnt_rgt)) { sigma = fmax(isleft, vl); sgndef = 1; }
    else { sigma = fmin(isright, vu); sgndef = -1; }
    if (mbl > 1) {
        double clwdth = W[wend] + Werr[wend] - W[wbgn] - Werr[wbgn];
        double avgap = fabs(clwdth / (double)(wend - wbgn));
        if (sgndef > 0) { tau = fmax(Wgap[wbgn], avgap); tau *= 0.5; tau = fmax(tau, Werr[wbgn]); }
        else { tau = fmax(Wgap[wend - 1], avgap); tau *= 0.5; tau = fmax(tau, Werr[wend]); }
    } else tau = Werr[wbgn];

    double *wd = (double *)malloc(sizeof(double) * (size_t)(2 * nbl));
    double *wl = wd + nbl;
    int ok = 0;
    for (int jtry = 0; jtry < MAX_TRY_ROOT; jtry++) {
        double Dpivot = D[ibgn] - sigma;
        wd[0] = Dpivot;
        double Dmax = fabs(wd[0]);
        int j = ibgn;
        for (int i = 0; i < nbl - 1; i++) {
            double invd = 1.0 / wd[i];
            double t = E[j] * invd;
            wl[i] = t;
            Dpivot = (D[j + 1] - sigma) - t * E[j];
            wd[i + 1] = Dpivot;
            Dmax = fmax(Dmax, fabs(Dpivot));
            j++;
        }
        int noREP = (Dmax > MAX_GROWTH * spdiam);
        if (noREP) {
            if (jtry == MAX_TRY_ROOT - 2) {
                if (sgndef == 1) sigma = gl - FUDGE_ROOT * spdiam * EPS * nbl - FUDGE_ROOT * 2.0 * pivmin;
                else sigma = gu + FUDGE_ROOT * spdiam * EPS * nbl + FUDGE_ROOT * 2.0 * pivmin;
            } else if (jtry == MAX_TRY_ROOT - 1) {
                free(wd);
                return 2;
            } else {
                sigma -= sgndef * tau;
                tau *= 2.0;
            }
        } else { ok = 1; break; }
    }
    if (!ok) { free(wd); return 2; }
    E[iend] = sigma;
    memcpy(D + ibgn, wd, sizeof(double) * (size_t)nbl);
    memcpy(E + ibgn, wl, sizeof(double) * (size_t)(nbl - 1));
    if (mbl > 1) {
        int iseed[4] = {1, 1, 1, 1};
        double *rv = (double *)malloc(sizeof(double) * (size_t)(2 * nbl));
        oracle_dlarnv2(iseed, 2 * nbl, rv);
        for (int i = 0; i < nbl - 1; i++) {
            D[i + ibgn] *= 1.0 + EPS * RAND_FACTOR * rv[i];
            E[i + ibgn] *= 1.0 + EPS * RAND_FACTOR * rv[i + nbl];
        }
        D[iend] *= 1.0 + EPS * RAND_FACTOR * rv[2 * nbl - 1];
        free(rv);
    }
    free(wd);
    *sgndefp = sgndef;
    *sigmap = sigma;
    *spdiamp = spdiam;
    return 0;
}

/* ==============================================================================
 * Eigenvector stage: mrrr_vec (mrrr_vec.c:75-445) with process_cluster.c and
 * process_singleton.c, on one thread.  Queues are FIFO as in queue.c; with one
 * thread the R-queue stays empty (process_cluster.c:296-299 gives own_part >=
 * cl_size), so refinement is always done in place.
 * ============================================================================ */
typedef struct {
    double *D, *L, *DL, *DLL;
} o_rrr;

typedef struct {
    int begin, end, depth, bl_begin, bl_end, bl_W_begin, bl_W_end;
    double bl_spdiam, lgap;
    o_rrr *RRR;
} o_task;

typedef struct {
    o_task **a;
    int head, tail, cap;
} o_queue;

static void q_push(o_queue *q, o_task *t)
{
    if (q->tail == q->cap) {
        q->cap = q->cap ? 2 * q->cap : 64;
        q->a = (o_task **)realloc(q->a, sizeof(o_task *) * (size_t)q->cap);
    }
    q->a[q->tail++] = t;
}
static o_task *q_pop(o_queue *q) { return (q->head < q->tail) ? q->a[q->head++] : NULL; }

typedef struct {
    o_state *st;
    double *Z;
    int ldz;
    int *Zsupp;
    o_queue sq, cq;
    int num_left;
    o_rrr **rrrs;       /* every representation created, freed at the end */
    int nrrr, caprrr;
    double *work;       /* 4n */
    int n;
} o_vec;

static o_rrr *new_rrr(o_vec *v, int size, int alloc_dl)
{
    o_rrr *r = (o_rrr *)calloc(1, sizeof(o_rrr));
    r->D = (double *)malloc(sizeof(double) * (size_t)size);
    r->L = (double *)malloc(sizeof(double) * (size_t)size);
    if (alloc_dl) {
        r->DL = (double *)malloc(sizeof(double) * (size_t)size);
        r->DLL = (double *)malloc(sizeof(double) * (size_t)size);
    }
    if (v->nrrr == v->caprrr) {
        v->caprrr = v->caprrr ? 2 * v->caprrr : 64;
        v->rrrs = (o_rrr **)realloc(v->rrrs, sizeof(o_rrr *) * (size_t)v->caprrr);
    }
    v->rrrs[v->nrrr++] = r;
    return r;
}

static o_task *new_task(int begin, int end, int depth, int bl_begin, int bl_end,
                        int bl_W_begin, int bl_W_end, double bl_spdiam,
                        double lgap, o_rrr *R)
{
    o_task *t = (o_task *)malloc(sizeof(o_task));
    t->begin = begin; t->end = end; t->depth = depth;
    t->bl_begin = bl_begin; t->bl_end = bl_end;
    t->bl_W_begin = bl_W_begin; t->bl_W_end = bl_W_end;
    t->bl_spdiam = bl_spdiam; t->lgap = lgap; t->RRR = R;
    return t;
}

/* PMR_process_s_task, process_singleton.c:65-324 */
static void o_process_s_task(o_vec *v, o_task *sng)
{
    o_state *st = v->st;
    int begin = sng->begin, end = sng->end, size = end - begin + 1;
    int bl_begin = sng->bl_begin, bl_end = sng->bl_end, bl_size = bl_end - bl_begin + 1;
    double bl_spdiam = sng->bl_spdiam;
    o_rrr *R = sng->RRR;
    double *D = R->D, *L = R->L, *DL = R->DL, *DLL = R->DLL;
    int n = st->n, ldz = v->ldz;
    double *W = st->W, *Werr = st->Werr, *Wgap = st->Wgap, *Wshifted = st->Wshifted;
    int *Windex = st->Windex, *Zindex = st->Zindex, *isuppZ = v->Zsupp;
    double *Z = v->Z, pivmin = st->pivmin;
    double tol = 4.0 * log((double)bl_size) * DBL_EPSILON;
    const double twoeps = 2 * DBL_EPSILON, RQtol = 2 * DBL_EPSILON;
    v->num_left -= size;
    for (int i = begin; i <= end; i++) {
        size_t zind = (size_t)Zindex[i];
        if (bl_size == 1) {
            memset(&Z[zind * ldz], 0, sizeof(double) * (size_t)n);
            Z[zind * ldz + bl_begin] = 1.0;
            isuppZ[2 * zind] = bl_begin + 1;
            isuppZ[2 * zind + 1] = bl_begin + 1;
            continue;
        }
        double lambda = Wshifted[i];
        double left = Wshifted[i] - Werr[i], right = Wshifted[i] + Werr[i];
        int i_local = Windex[i], r = 0;
        double lgap, rgap, gap, gaptol;
        if (i == bl_begin) lgap = DBL_EPSILON * fmax(fabs(left), fabs(right));
        else if (i == begin) lgap = sng->lgap;
        else lgap = Wgap[i - 1];
        if (i == bl_end) rgap = DBL_EPSILON * fmax(fabs(left), fabs(right));
        else rgap = Wgap[i];
        gap = fmin(lgap, rgap);
        gaptol = (i == bl_begin || i == bl_end) ? 0.0 : gap * DBL_EPSILON;
        int i_supmn = bl_size, i_supmx = 1;
        double savedgap = Wgap[i];
        Wgap[i] = gap;
        int usedBS = 0, usedRQ = 0, needBS = 0, wantNC, negcount = 0;
        double bstres = 0, bstw = 0, residual = 0, RQcorr = 0, ztz, mingma, norminv = 0;
        memset(&Z[zind * ldz], 0, sizeof(double) * (size_t)n);
        for (int k = 1; k <= MAXITER_S; k++) {
            if (needBS) {
                ++o_fallbacks;
                usedBS = 1;
                int itmp = r, offset = Windex[i] - 1;
                double tmp = Wgap[i];
                Wgap[i] = 0.0;
                oracle_dlarrb(bl_size, D, DLL, i_local, i_local, 0.0, twoeps, offset,
                              &Wshifted[i], &Wgap[i], &Werr[i], pivmin, bl_spdiam, itmp);
                Wgap[i] = tmp;
                lambda = Wshifted[i];
                r = 0;
            }
            wantNC = usedBS ? 0 : 1;
            oracle_dlar1v(bl_size, 1, bl_size, lambda, D, L, DL, DLL, pivmin, gaptol,
                          &Z[zind * ldz + bl_begin], wantNC, &negcount, &ztz, &mingma,
                          &r, &isuppZ[2 * zind], &norminv, &residual, &RQcorr, v->work);
            if (k == 1 || residual < bstres) { bstres = residual; bstw = lambda; }
            i_supmn = imin(i_supmn, isuppZ[2 * zind]);
            i_supmx = imax(i_supmx, isuppZ[2 * zind + 1]);
            if (!usedBS && residual > tol * gap && fabs(RQcorr) > RQtol * fabs(lambda)) {
                int sgndef = (i_local <= negcount) ? -1 : 1;
                if (RQcorr * sgndef >= 0.0 && lambda + RQcorr <= right && lambda + RQcorr >= left) {
                    usedRQ = 1;
                    if (sgndef == 1) left = lambda;
                    else right = lambda;
                    Wshifted[i] = 0.5 * (left + right);
                    lambda += RQcorr;
                } else needBS = 1;
                if (right - left < RQtol * fabs(lambda)) usedBS = 1;
                else if (k == MAXITER_S - 1) needBS = 1;
            } else break;
        }
        if (usedRQ && usedBS && bstres <= residual) {
            lambda = bstw;
            oracle_dlar1v(bl_size, 1, bl_size, lambda, D, L, DL, DLL, pivmin, gaptol,
                          &Z[zind * ldz + bl_begin], wantNC, &negcount, &ztz, &mingma,
                          &r, &isuppZ[2 * zind], &norminv, &residual, &RQcorr, v->work);
        }
        Wshifted[i] = lambda;
        isuppZ[2 * zind] += bl_begin;
        isuppZ[2 * zind + 1] += bl_begin;
        int i_Zfrom = isuppZ[2 * zind] - 1, i_Zto = isuppZ[2 * zind + 1] - 1;
        i_supmn += bl_begin - 1;
        i_supmx += bl_begin - 1;
        for (int k = i_supmn; k < i_Zfrom; k++) Z[k + zind * ldz] = 0.0;
        for (int k = i_Zto + 1; k <= i_supmx; k++) Z[k + zind * ldz] = 0.0;
        for (int k = i_Zfrom; k <= i_Zto; k++) Z[k + zind * ldz] *= norminv;
        double sigma = L[bl_size - 1];
        W[i] = lambda + sigma;
        if (i < end) Wgap[i] = fmax(savedgap, W[i + 1] - Werr[i + 1] - W[i] - Werr[i]);
    }
    free(sng);
}

/* PMR_create_subtasks, process_cluster.c:438-629 (one thread: no immediate
 * processing difference, S-tasks with depth > 0 and cl_size < STASK_NOENQUEUE are
 * processed at once as in :530-535) */
static void o_create_subtasks(o_vec *v, o_task *cl, o_rrr *R)
{
    o_state *st = v->st;
    int cl_begin = cl->begin, cl_end = cl->end, cl_size = cl_end - cl_begin + 1;
    int depth = cl->depth, bl_begin = cl->bl_begin, bl_end = cl->bl_end;
    int bl_size = bl_end - bl_begin + 1;
    double bl_spdiam = cl->bl_spdiam;
    double *Wgap = st->Wgap, *Wshifted = st->Wshifted;
    int max_size = (int)fmax(1, v->num_left / 4);
    int task_inserted = 1, sn_first = -1, sn_last = -1, sn_size = -1;
    const double avggap_factor = 2.0;
    int new_first = cl_begin;
    for (int i = cl_begin; i <= cl_end; i++) {
        int new_last;
        if (i == cl_end) new_last = i;
        else if (Wgap[i] >= MIN_RELGAP * fabs(Wshifted[i])) new_last = i;
        else continue;
        int new_size = new_last - new_first + 1;
        if (new_size == 1) {
            if (new_first == cl_begin || task_inserted) {
                sn_first = new_first; sn_last = new_first; sn_size = 1;
            } else { sn_last++; sn_size++; }
            if (i == cl_end || sn_size >= max_size ||
                Wgap[i + 1] < MIN_RELGAP * fabs(Wshifted[i + 1])) {
                double lgap = (sn_first == cl_begin) ? cl->lgap : Wgap[sn_first - 1];
                o_task *t = new_task(sn_first, sn_last, depth + 1, bl_begin, bl_end,
                                     cl->bl_W_begin, cl->bl_W_end, bl_spdiam, lgap, R);
                if (depth > 0 && cl_size < STASK_NOENQUEUE) o_process_s_task(v, t);
                else q_push(&v->sq, t);
                task_inserted = 1;
            } else task_inserted = 0;
        } else {
            if (depth == 0 && new_size > 3) {
                int cl_first = new_first, cl_last;
                double avggap = bl_spdiam / (bl_size - 1);
                for (int k = new_first + 1; k < new_last; k++) {
                    if (k == new_last - 1) cl_last = new_last;
                    else if (k != cl_first && Wgap[k] > avggap_factor * avggap) cl_last = k;
                    else continue;
                    double lgap = (cl_first == cl_begin) ? cl->lgap : Wgap[cl_first - 1];
                    q_push(&v->cq, new_task(cl_first, cl_last, depth + 1, bl_begin, bl_end,
                                            cl->bl_W_begin, cl->bl_W_end, bl_spdiam, lgap, R));
                    cl_first = k + 1;
                }
            } else {
                double lgap = (new_first == cl_begin) ? cl->lgap : Wgap[new_first - 1];
                o_rrr *P = R;
                if (depth > 0) { /* parent RRR copy (the reference stashes it in Z, :597-602) */
                    P = new_rrr(v, bl_size, 0);
                    memcpy(P->D, R->D, sizeof(double) * (size_t)bl_size);
                    memcpy(P->L, R->L, sizeof(double) * (size_t)bl_size);
                }
                q_push(&v->cq, new_task(new_first, new_last, depth + 1, bl_begin, bl_end,
                                        cl->bl_W_begin, cl->bl_W_end, bl_spdiam, lgap, P));
            }
            task_inserted = 1;
        }
        new_first = i + 1;
    }
    free(cl);
}

/* PMR_process_c_task = compute_new_rrr + refine_eigvals, process_cluster.c:75-433 */
static void o_process_c_task(o_vec *v, o_task *cl)
{
    o_state *st = v->st;
    int cl_begin = cl->begin, cl_end = cl->end, cl_size = cl_end - cl_begin + 1;
    int depth = cl->depth, bl_begin = cl->bl_begin, bl_end = cl->bl_end;
    int bl_size = bl_end - bl_begin + 1;
    double bl_spdiam = cl->bl_spdiam;
    double *W = st->W, *Werr = st->Werr, *Wgap = st->Wgap, *Wshifted = st->Wshifted;
    int *Windex = st->Windex;
    double pivmin = st->pivmin, RQtol = st->RQtol;
    o_rrr *Rp = cl->RRR, *R;
    if (depth == 0) {
        R = new_rrr(v, 0, 0);
        free(R->D); free(R->L);
        R->D = Rp->D; R->L = Rp->L;        /* root RRR lives in D, E */
        R->DL = (double *)malloc(sizeof(double) * (size_t)bl_size);
        R->DLL = (double *)malloc(sizeof(double) * (size_t)bl_size);
    } else {
        R = new_rrr(v, bl_size, 1);
        double *DLp = (double *)malloc(sizeof(double) * (size_t)bl_size);
        double *DLLp = (double *)malloc(sizeof(double) * (size_t)bl_size);
        for (int i = 0; i < bl_size - 1; i++) {
            double tmp = Rp->D[i] * Rp->L[i];
            DLp[i] = tmp;
            DLLp[i] = tmp * Rp->L[i];
        }
        for (int k = 0; k < 2; k++) { /* :168-195 */
            int p;
            double savegap;
            if (k == 0) { p = Windex[cl_begin]; savegap = Wgap[cl_begin]; Wgap[cl_begin] = 0.0; }
            else { p = Windex[cl_end]; savegap = Wgap[cl_end]; Wgap[cl_end] = 0.0; }
            int offset = Windex[cl_begin] - 1;
            oracle_dlarrb(bl_size, Rp->D, DLLp, p, p, RQtol, RQtol, offset,
                          &Wshifted[cl_begin], &Wgap[cl_begin], &Werr[cl_begin],
                          pivmin, bl_spdiam, bl_size);
            if (k == 0)
                Wgap[cl_begin] = fmax(0, (Wshifted[cl_begin + 1] - Werr[cl_begin + 1]) -
                                         (Wshifted[cl_begin] + Werr[cl_begin]));
            else Wgap[cl_end] = savegap;
        }
        double lgap = cl->lgap, rgap = Wgap[cl_end], tau;
        oracle_dlarrf(bl_size, Rp->D, Rp->L, DLp, 1, cl_size, &Wshifted[cl_begin],
                      &Wgap[cl_begin], &Werr[cl_begin], bl_spdiam, lgap, rgap, pivmin,
                      &tau, R->D, R->L, v->work);
        R->L[bl_size - 1] = Rp->L[bl_size - 1] + tau;
        for (int k = cl_begin; k <= cl_end; k++) {
            double fudge = 3.0 * DBL_EPSILON * fabs(Wshifted[k]);
            Wshifted[k] -= tau;
            fudge += 4.0 * DBL_EPSILON * fabs(Wshifted[k]);
            Werr[k] += fudge;
        }
        free(DLp);
        free(DLLp);
    }
    for (int i = 0; i < bl_size - 1; i++) {
        double tmp = R->D[i] * R->L[i];
        R->DL[i] = tmp;
        R->DLL[i] = tmp * R->L[i];
    }
    /* refine_eigvals, :240-433 */
    int mbl = cl->bl_W_end - cl->bl_W_begin + 1;
    if (depth == 0) {
        double gl = st->gersch[2 * bl_begin], gu = st->gersch[2 * bl_begin + 1];
        for (int i = bl_begin + 1; i < bl_end; i++) {
            gl = fmin(gl, st->gersch[2 * i]);
            gu = fmax(gu, st->gersch[2 * i + 1]);
        }
        cl->bl_spdiam = gu - gl;
        double sigma = R->L[bl_size - 1];
        memcpy(&Wshifted[cl_begin], &W[cl_begin], sizeof(double) * (size_t)mbl);
        for (int i = 0; i < mbl; i++) W[cl_begin + i] += sigma;
    } else {
        int p = Windex[cl_begin], q = Windex[cl_end], offset = Windex[cl_begin] - 1;
        double savegap = 0;
        if (p == q) { savegap = Wgap[cl_begin]; Wgap[cl_begin] = 0.0; }
        oracle_dlarrb(bl_size, R->D, R->DLL, p, q, st->rtol1, st->rtol2, offset,
                      &Wshifted[cl_begin], &Wgap[cl_begin], &Werr[cl_begin], pivmin,
                      bl_spdiam, bl_size);
        if (p == q) Wgap[cl_begin] = savegap;
        double sigma = R->L[bl_size - 1];
        for (int j = cl_begin; j <= cl_end; j++) W[j] = Wshifted[j] + sigma;
    }
    (void)cl_size;
    o_create_subtasks(v, cl, R);
}

static int cmp_zindex_key(const void *a, const void *b);
typedef struct { double lambda; int ilocal, iblock, ind; } o_sort;

/* mrrr_vec.c:75-143 with init_zindex (:380-416) and init_workQ (:304-374) */
static void o_mrrr_vec(o_state *st, double vl, double *Z, int ldz, int *Zsupp)
{
    int n = st->n, m = st->m;
    o_vec v;
    memset(&v, 0, sizeof(v));
    v.st = st; v.Z = Z; v.ldz = ldz; v.Zsupp = Zsupp; v.n = n;
    v.work = (double *)malloc(sizeof(double) * (size_t)(4 * n + 4));
    st->Wshifted = (double *)malloc(sizeof(double) * (size_t)n);
    double *L = st->E;
    o_sort *arr = (o_sort *)malloc(sizeof(o_sort) * (size_t)(m > 0 ? m : 1));
    for (int i = 0; i < m; i++) {
        int iblk = st->iblock[i];
        int ishift = st->isplit[iblk - 1] - 1;
        arr[i].lambda = st->W[i] + L[ishift];
        arr[i].ilocal = st->Windex[i];
        arr[i].iblock = iblk;
        arr[i].ind = i;
    }
    qsort(arr, (size_t)m, sizeof(o_sort), cmp_zindex_key);
    for (int i = 0; i < m; i++) st->Zindex[arr[i].ind] = i;
    free(arr);
    v.num_left = m;

    int begin = 0, Wbegin = 0;
    for (int j = 0; j < st->nsplit; j++) {
        int end = st->isplit[j] - 1;
        double sigma = L[end];
        int Wend = Wbegin - 1;
        while (Wend < m - 1 && st->iblock[Wend + 1] == j + 1) Wend++;
        if (Wend < Wbegin) { begin = end + 1; continue; }
        int nbl = end - begin + 1;
        o_rrr *R = new_rrr(&v, 0, 0);
        free(R->D); free(R->L);
        R->D = &st->D[begin]; R->L = &L[begin];
        if (nbl == 1) {
            q_push(&v.sq, new_task(Wbegin, Wbegin, 1, begin, end, Wbegin, Wend, 0, 0, R));
            begin = end + 1;
            Wbegin = Wend + 1;
            continue;
        }
        double lgap = fmax(0.0, (st->W[Wbegin] + sigma) - st->Werr[Wbegin] - vl);
        q_push(&v.cq, new_task(Wbegin, Wend, 0, begin, end, Wbegin, Wend, 0, lgap, R));
        begin = end + 1;
        Wbegin = Wend + 1;
    }
    for (;;) { /* empty_workQ, mrrr_vec.c:259-291 (S before C) */
        o_task *t = q_pop(&v.sq);
        if (t) { o_process_s_task(&v, t); continue; }
        t = q_pop(&v.cq);
        if (t) { o_process_c_task(&v, t); continue; }
        break;
    }
    for (int i = 0; i < v.nrrr; i++) {
        o_rrr *r = v.rrrs[i];
        /* root representations point into D, E and are not owned */
        int owned = !(r->D >= st->D && r->D < st->D + n);
        if (owned) { free(r->D); free(r->L); }
        free(r->DL); free(r->DLL);
        free(r);
    }
    free(v.rrrs);
    free(v.sq.a);
    free(v.cq.a);
    free(v.work);
    free(st->Wshifted);
    st->Wshifted = NULL;
}

static int cmp_zindex_key(const void *a1, const void *a2)
{
    const o_sort *x = (const o_sort *)a1, *y = (const o_sort *)a2;
    if (x->iblock == y->iblock) return x->ilocal - y->ilocal;
    if (x->lambda < y->lambda) return -1;
    if (x->lambda > y->lambda) return 1;
    return (x->ilocal < y->ilocal) ? -1 : 1;
}

static int cmpa(const void *a1, const void *a2)
{
    double x = *(const double *)a1, y = *(const double *)a2;
    return (x < y) ? -1 : 1;
}

typedef struct { double lambda; int ilocal; int ind; } o_sort2;
static int cmpb(const void *a1, const void *a2)
{
    const o_sort2 *x = (const o_sort2 *)a1, *y = (const o_sort2 *)a2;
    return (x->ind < y->ind) ? -1 : 1;
}

/* dstemr's n = 1 and n = 2 cases, odstmr.f:357-430 */
/* ==============================================================================
 * mrrr(): mrrr.c:87-296 (with the scaling of :325-386, rel-acc refinement of
 * :452-520 and the eigenpair sort of :525-588).
 * ============================================================================ */
int oracle_mrrr(char *jobz, char *range, int *np, double *D, double *E,
                double *vlp, double *vup, int *ilp, int *iup, int *tryracp,
                int *mp, double *W, double *Z, int *ldz, int *Zsupp)
{
    int n = *np;
    int onlyW = (jobz[0] == 'N' || jobz[0] == 'n');
    int wantZ = (jobz[0] == 'V' || jobz[0] == 'v');
    int cntval = (jobz[0] == 'C' || jobz[0] == 'c');
    int alleig = (range[0] == 'A' || range[0] == 'a');
    int valeig = (range[0] == 'V' || range[0] == 'v');
    int indeig = (range[0] == 'I' || range[0] == 'i');
    if (!(onlyW || wantZ || cntval)) return 1;
    if (!(alleig || valeig || indeig)) return 1;
    if (n <= 0) return 1;
    if (valeig) { if (*vup <= *vlp) return 1; }
    else if (indeig) { if (*ilp < 1 || *ilp > n || *iup < *ilp || *iup > n) return 1; }
    if (cntval) {
        if (alleig) { *mp = n; return 0; }
        if (indeig) { *mp = *iup - *ilp + 1; return 0; }
    }
    if (n < MR3_DSTEMR_BELOW) { /* handle_small_cases, mrrr.c:149-153, :394-441 */
        if (cntval) {
            double cnt = 0.0;
            int ld1 = *ldz > 0 ? *ldz : 1;
            int rc = oracle_dstemr('V', 'V', n, D, E, *vlp, *vup, *ilp, *iup, mp, W, &cnt,
                                   ld1 < n ? n : ld1, -1, Zsupp, tryracp);
            *mp = (int)cnt;
            return rc == 0 ? 0 : 2;
        }
        int rc = oracle_dstemr(wantZ ? 'V' : 'N', alleig ? 'A' : (valeig ? 'V' : 'I'), n, D, E,
                               *vlp, *vup, *ilp, *iup, mp, W, Z, *ldz, n, Zsupp, tryracp);
        return rc == 0 ? 0 : 2;
    }

    char jz = wantZ ? 'V' : (onlyW ? 'N' : 'C');
    char rg = alleig ? 'A' : (valeig ? 'V' : 'I');
    o_state st;
    memset(&st, 0, sizeof(st));
    st.n = n; st.D = D; st.E = E;
    st.E2 = (double *)malloc(sizeof(double) * (size_t)n);
    st.isplit = (int *)malloc(sizeof(int) * (size_t)n);
    st.W = W;
    st.Werr = (double *)malloc(sizeof(double) * (size_t)n);
    st.Wgap = (double *)malloc(sizeof(double) * (size_t)n);
    st.gersch = (double *)malloc(sizeof(double) * (size_t)(2 * n));
    st.iblock = (int *)calloc((size_t)n, sizeof(int));
    st.Windex = (int *)malloc(sizeof(int) * (size_t)n);
    st.Zindex = (int *)malloc(sizeof(int) * (size_t)n);

    /* scale_matrix, mrrr.c:325-365 */
    double scale = 1.0, smlnum = DBL_MIN / DBL_EPSILON, bignum = 1.0 / smlnum;
    double rmin = sqrt(smlnum), rmax = fmin(sqrt(bignum), 1.0 / sqrt(sqrt(DBL_MIN)));
    double T_norm = 0.0;
    for (int i = 0; i < n; i++) { double a = fabs(D[i]); if (a > T_norm || isnan(a)) T_norm = a; }
    for (int i = 0; i < n - 1; i++) { double a = fabs(E[i]); if (a > T_norm || isnan(a)) T_norm = a; }
    if (T_norm > 0 && T_norm < rmin) scale = rmin / T_norm;
    else if (T_norm > rmax) scale = rmax / T_norm;
    if (scale != 1.0) {
        for (int i = 0; i < n; i++) D[i] *= scale;
        for (int i = 0; i < n - 1; i++) E[i] *= scale;
        if (valeig) { *vlp *= scale; *vup *= scale; }
    }
    int info;
    double *Dcopy = NULL, *E2copy = NULL;
    if (*tryracp) info = o_dlarrr(n, D, E);
    else info = -1;
    if (info == 0) {
        st.split = DBL_EPSILON;
        Dcopy = (double *)malloc(sizeof(double) * (size_t)n);
        memcpy(Dcopy, D, sizeof(double) * (size_t)n);
        E2copy = (double *)malloc(sizeof(double) * (size_t)n);
        for (int i = 0; i < n - 1; i++) E2copy[i] = E[i] * E[i];
        E2copy[n - 1] = 0.0;
    } else {
        st.split = -DBL_EPSILON;
        *tryracp = 0;
    }
    if (!wantZ) { st.rtol1 = 4.0 * DBL_EPSILON; st.rtol2 = 4.0 * DBL_EPSILON; }
    else {
        st.rtol1 = fmin(sqrt(DBL_EPSILON), 1e-2 * MIN_RELGAP);
        st.rtol2 = fmax(fmin(st.rtol1 * 5.0E-3, MIN_RELGAP * 5.0E-6), 4.0 * DBL_EPSILON);
    }
    st.bsrtol = fmin(MIN_RELGAP * 1.0E-2, sqrt(DBL_EPSILON));
    st.RQtol = 2.0 * DBL_EPSILON;

    int rc = o_mrrr_val(jz, rg, &st, vlp, vup, ilp, iup);
    *mp = st.m;
    if (rc != 0 || cntval) goto out;
    if (wantZ) {
        o_mrrr_vec(&st, *vlp, Z, *ldz, Zsupp);
    } else {
        for (int i = 0; i < *mp; i++) {
            int iblk = st.iblock[i];
            int ishift = st.isplit[iblk - 1] - 1;
            W[i] += E[ishift];
        }
    }
    if (*tryracp) { /* refine_to_highrac, mrrr.c:452-520 */
        int begin = 0, Wbegin = 0;
        for (int j = 0; j < st.nsplit; j++) {
            int end = st.isplit[j] - 1, nbl = end - begin + 1, mbl = 0;
            int ilow = n, iupp = 1;
            for (int i = Wbegin; i < *mp; i++) {
                if (st.iblock[i] == j + 1) {
                    mbl++;
                    ilow = imin(ilow, st.Windex[i]);
                    iupp = imax(iupp, st.Windex[i]);
                } else break;
            }
            if (mbl == 0) { begin = end + 1; continue; }
            int Wend = Wbegin + mbl - 1;
            o_dlarrj(nbl, &Dcopy[begin], &E2copy[begin], ilow, iupp, 4 * DBL_EPSILON,
                     ilow - 1, &W[Wbegin], &st.Werr[Wbegin], st.pivmin, st.spdiam);
            begin = end + 1;
            Wbegin = Wend + 1;
        }
    }
    if (scale != 1.0) {
        double inv = 1.0 / scale;
        *vlp *= inv; *vup *= inv;
        for (int i = 0; i < *mp; i++) W[i] *= inv;
    }
    if (wantZ) { /* sort_eigenpairs, mrrr.c:525-588 */
        int m = *mp;
        o_sort2 *arr = (o_sort2 *)malloc(sizeof(o_sort2) * (size_t)(m > 0 ? m : 1));
        for (int i = 0; i < m; i++) {
            arr[i].lambda = W[i]; arr[i].ilocal = st.Windex[i]; arr[i].ind = st.Zindex[i];
        }
        qsort(arr, (size_t)m, sizeof(o_sort2), cmpb);
        for (int i = 0; i < m; i++) { W[i] = arr[i].lambda; st.Windex[i] = arr[i].ilocal; }
        free(arr);
        double *tmpc = (double *)malloc(sizeof(double) * (size_t)n);
        int sorted = 0;
        while (!sorted) {
            sorted = 1;
            for (int i = 0; i < m - 1; i++) {
                if (W[i] > W[i + 1]) {
                    sorted = 0;
                    double t = W[i]; W[i] = W[i + 1]; W[i + 1] = t;
                    int t1 = Zsupp[2 * i]; Zsupp[2 * i] = Zsupp[2 * (i + 1)]; Zsupp[2 * (i + 1)] = t1;
                    int t2 = Zsupp[2 * i + 1]; Zsupp[2 * i + 1] = Zsupp[2 * (i + 1) + 1]; Zsupp[2 * (i + 1) + 1] = t2;
                    memcpy(tmpc, &Z[(size_t)i * *ldz], sizeof(double) * (size_t)n);
                    memcpy(&Z[(size_t)i * *ldz], &Z[(size_t)(i + 1) * *ldz], sizeof(double) * (size_t)n);
                    memcpy(&Z[(size_t)(i + 1) * *ldz], tmpc, sizeof(double) * (size_t)n);
                }
            }
        }
        free(tmpc);
    } else {
        qsort(W, (size_t)*mp, sizeof(double), cmpa);
    }
out:
    free(st.E2); free(st.isplit); free(st.Werr); free(st.Wgap); free(st.gersch);
    free(st.iblock); free(st.Windex); free(st.Zindex);
    free(Dcopy); free(E2copy);
    return rc == 0 ? 0 : 2;
}
