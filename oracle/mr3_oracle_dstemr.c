/*
 * mr3_oracle_dstemr.c -- CPU oracle, second part: the sequential route of mrrr().
 *
 * TEST INFRASTRUCTURE ONLY (see mr3_oracle.h).  The reference's mrrr() hands every
 * matrix with n < 500 (and every matrix when it runs on one thread) to LAPACK's
 * dstemr (SRC/mrrr.c:149-153, :394-441), vendored as SRC/LAPACK/odstmr.f.  This
 * file restates that route:
 *   dstemr  odstmr.f:233-647   scaling, dlarrr, dlarre, dlarrv, dlarrj, sort
 *   dlarre  odrre.f:296-891    root representations; dqds (range A) or bisection
 *   dlasq2  odsq2.f (+ odsq3.f, odsq4.f, odsq5.f, odsq6.f)   dqds
 *   dlarrk  odrrk.f:95-167     one eigenvalue of T by bisection
 *   dlarrv  odrrv.f:150-896    the representation tree, breadth first
 * Fortran's left-to-right evaluation order is kept in every expression, and the
 * file is compiled with -ffp-contract=off, so results round as the reference's do.
 */
#include "mr3_oracle.h"
#include "mr3_oracle_int.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define EPS    DBL_EPSILON
#define SAFMIN DBL_MIN

static inline int    mini(int a, int b) { return a < b ? a : b; }
static inline int    maxi(int a, int b) { return a > b ? a : b; }
/* Fortran MIN/MAX of two reals (no NaN reaches them on this route) */
static inline double fmn(double a, double b) { return b < a ? b : a; }
static inline double fmx(double a, double b) { return b > a ? b : a; }

/* ------------------------------------------------------------------------------
 * dlarrk: eigenvalue IW of T by bisection to relative width RELTOL, odrrk.f:95-167.
 * ---------------------------------------------------------------------------- */
static void o_dlarrk(int n, int iw, double gl, double gu, const double *d,
                     const double *e2, double pivmin, double reltol, double *w,
                     double *werr)
{
    const double FUDGE = 2.0;
    double tnorm = fmx(fabs(gl), fabs(gu));
    double rtoli = reltol, atoli = FUDGE * 2.0 * pivmin;
    int itmax = (int)((log(tnorm + pivmin) - log(pivmin)) / log(2.0)) + 2;
    double left = gl - FUDGE * tnorm * EPS * n - FUDGE * 2.0 * pivmin;
    double right = gu + FUDGE * tnorm * EPS * n + FUDGE * 2.0 * pivmin;
    int it = 0;
    for (;;) {
        double tmp1 = fabs(right - left);
        double tmp2 = fmx(fabs(right), fabs(left));
        if (tmp1 < fmx(fmx(atoli, pivmin), rtoli * tmp2)) break;
        if (it > itmax) break;
        it++;
        double mid = 0.5 * (left + right);
        int negcnt = 0;
        tmp1 = d[0] - mid;
        if (fabs(tmp1) < pivmin) tmp1 = -pivmin;
        if (tmp1 <= 0.0) negcnt++;
        for (int i = 2; i <= n; i++) {
            tmp1 = d[i - 1] - e2[i - 2] / tmp1 - mid;
            if (fabs(tmp1) < pivmin) tmp1 = -pivmin;
            if (tmp1 <= 0.0) negcnt++;
        }
        if (negcnt >= iw) right = mid;
        else left = mid;
    }
    *w = 0.5 * (left + right);
    *werr = 0.5 * fabs(right - left);
}

/* ------------------------------------------------------------------------------
 * dqds: dlasq2 and its helpers dlasq3-6.  Z is the 1-based qd array of 4n entries.
 * ---------------------------------------------------------------------------- */
#define Z(i) z[(i) - 1]

/* dlasq4: the shift, odsq4.f.  TAU is left as it was on the early returns. */
static void o_dlasq4(int i0, int n0, const double *z, int pp, int n0in, double dmin,
                     double dmin1, double dmin2, double dn, double dn1, double dn2,
                     double *tau, int *ttype, double *g)
{
    const double CNST1 = 0.5630, CNST2 = 1.010, CNST3 = 1.050;
    const double QURTR = 0.250, THIRD = 0.3330, HALF = 0.50;
    double a2, b1, b2, gam, gap1, gap2, s = 0.0;
    int i4, nn, np;
    if (dmin <= 0.0) { *tau = -dmin; *ttype = -1; return; }
    nn = 4 * n0 + pp;
    if (n0in == n0) {
        if (dmin == dn || dmin == dn1) {
            b1 = sqrt(Z(nn - 3)) * sqrt(Z(nn - 5));
            b2 = sqrt(Z(nn - 7)) * sqrt(Z(nn - 9));
            a2 = Z(nn - 7) + Z(nn - 5);
            if (dmin == dn && dmin1 == dn1) {
                gap2 = dmin2 - a2 - dmin2 * QURTR;
                if (gap2 > 0.0 && gap2 > b2) gap1 = a2 - dn - (b2 / gap2) * b2;
                else gap1 = a2 - dn - (b1 + b2);
                if (gap1 > 0.0 && gap1 > b1) {
                    s = fmx(dn - (b1 / gap1) * b1, HALF * dmin);
                    *ttype = -2;
                } else {
                    s = 0.0;
                    if (dn > b1) s = dn - b1;
                    if (a2 > (b1 + b2)) s = fmn(s, a2 - (b1 + b2));
                    s = fmx(s, THIRD * dmin);
                    *ttype = -3;
                }
            } else {
                *ttype = -4;
                s = QURTR * dmin;
                if (dmin == dn) {
                    gam = dn;
                    a2 = 0.0;
                    if (Z(nn - 5) > Z(nn - 7)) return;
                    b2 = Z(nn - 5) / Z(nn - 7);
                    np = nn - 9;
                } else {
                    np = nn - 2 * pp;
                    b2 = Z(np - 2);
                    gam = dn1;
                    if (Z(np - 4) > Z(np - 2)) return;
                    a2 = Z(np - 4) / Z(np - 2);
                    if (Z(nn - 9) > Z(nn - 11)) return;
                    b2 = Z(nn - 9) / Z(nn - 11);
                    np = nn - 13;
                }
                a2 = a2 + b2;
                for (i4 = np; i4 >= 4 * i0 - 1 + pp; i4 -= 4) {
                    if (b2 == 0.0) break;
                    b1 = b2;
                    if (Z(i4) > Z(i4 - 2)) return;
                    b2 = b2 * (Z(i4) / Z(i4 - 2));
                    a2 = a2 + b2;
                    if (100.0 * fmx(b2, b1) < a2 || CNST1 < a2) break;
                }
                a2 = CNST3 * a2;
                if (a2 < CNST1) s = gam * (1.0 - sqrt(a2)) / (1.0 + a2);
            }
        } else if (dmin == dn2) {
            *ttype = -5;
            s = QURTR * dmin;
            np = nn - 2 * pp;
            b1 = Z(np - 2);
            b2 = Z(np - 6);
            gam = dn2;
            if (Z(np - 8) > b2 || Z(np - 4) > b1) return;
            a2 = (Z(np - 8) / b2) * (1.0 + Z(np - 4) / b1);
            if (n0 - i0 > 2) {
                b2 = Z(nn - 13) / Z(nn - 15);
                a2 = a2 + b2;
                for (i4 = nn - 17; i4 >= 4 * i0 - 1 + pp; i4 -= 4) {
                    if (b2 == 0.0) break;
                    b1 = b2;
                    if (Z(i4) > Z(i4 - 2)) return;
                    b2 = b2 * (Z(i4) / Z(i4 - 2));
                    a2 = a2 + b2;
                    if (100.0 * fmx(b2, b1) < a2 || CNST1 < a2) break;
                }
                a2 = CNST3 * a2;
            }
            if (a2 < CNST1) s = gam * (1.0 - sqrt(a2)) / (1.0 + a2);
        } else {
            if (*ttype == -6) *g = *g + THIRD * (1.0 - *g);
            else if (*ttype == -18) *g = QURTR * THIRD;
            else *g = QURTR;
            s = *g * dmin;
            *ttype = -6;
        }
    } else if (n0in == n0 + 1) {
        if (dmin1 == dn1 && dmin2 == dn2) {
            *ttype = -7;
            s = THIRD * dmin1;
            if (Z(nn - 5) > Z(nn - 7)) return;
            b1 = Z(nn - 5) / Z(nn - 7);
            b2 = b1;
            if (b2 != 0.0) {
                for (i4 = 4 * n0 - 9 + pp; i4 >= 4 * i0 - 1 + pp; i4 -= 4) {
                    a2 = b1;
                    if (Z(i4) > Z(i4 - 2)) return;
                    b1 = b1 * (Z(i4) / Z(i4 - 2));
                    b2 = b2 + b1;
                    if (100.0 * fmx(b1, a2) < b2) break;
                }
            }
            b2 = sqrt(CNST3 * b2);
            a2 = dmin1 / (1.0 + b2 * b2);
            gap2 = HALF * dmin2 - a2;
            if (gap2 > 0.0 && gap2 > b2 * a2) s = fmx(s, a2 * (1.0 - CNST2 * a2 * (b2 / gap2) * b2));
            else { s = fmx(s, a2 * (1.0 - CNST2 * b2)); *ttype = -8; }
        } else {
            s = QURTR * dmin1;
            if (dmin1 == dn1) s = HALF * dmin1;
            *ttype = -9;
        }
    } else if (n0in == n0 + 2) {
        if (dmin2 == dn2 && 2.0 * Z(nn - 5) < Z(nn - 7)) {
            *ttype = -10;
            s = THIRD * dmin2;
            if (Z(nn - 5) > Z(nn - 7)) return;
            b1 = Z(nn - 5) / Z(nn - 7);
            b2 = b1;
            if (b2 != 0.0) {
                for (i4 = 4 * n0 - 9 + pp; i4 >= 4 * i0 - 1 + pp; i4 -= 4) {
                    if (Z(i4) > Z(i4 - 2)) return;
                    b1 = b1 * (Z(i4) / Z(i4 - 2));
                    b2 = b2 + b1;
                    if (100.0 * b1 < b2) break;
                }
            }
            b2 = sqrt(CNST3 * b2);
            a2 = dmin2 / (1.0 + b2 * b2);
            gap2 = Z(nn - 7) + Z(nn - 9) - sqrt(Z(nn - 11)) * sqrt(Z(nn - 9)) - a2;
            if (gap2 > 0.0 && gap2 > b2 * a2) s = fmx(s, a2 * (1.0 - CNST2 * a2 * (b2 / gap2) * b2));
            else s = fmx(s, a2 * (1.0 - CNST2 * b2));
        } else {
            s = QURTR * dmin2;
            *ttype = -11;
        }
    } else if (n0in > n0 + 2) {
        s = 0.0;
        *ttype = -12;
    }
    *tau = s;
}

/* dlasq5: one dqds transform with shift TAU, IEEE form, odsq5.f. */
static void o_dlasq5(int i0, int n0, double *z, int pp, double tau, double *dmin,
                     double *dmin1, double *dmin2, double *dn, double *dnm1,
                     double *dnm2)
{
    int j4, j4p2;
    double d, emin, temp;
    if ((n0 - i0 - 1) <= 0) return;
    j4 = 4 * i0 + pp - 3;
    emin = Z(j4 + 4);
    d = Z(j4) - tau;
    *dmin = d;
    *dmin1 = -Z(j4);
    if (pp == 0) {
        for (j4 = 4 * i0; j4 <= 4 * (n0 - 3); j4 += 4) {
            Z(j4 - 2) = d + Z(j4 - 1);
            temp = Z(j4 + 1) / Z(j4 - 2);
            d = d * temp - tau;
            *dmin = fmn(*dmin, d);
            Z(j4) = Z(j4 - 1) * temp;
            emin = fmn(Z(j4), emin);
        }
    } else {
        for (j4 = 4 * i0; j4 <= 4 * (n0 - 3); j4 += 4) {
            Z(j4 - 3) = d + Z(j4);
            temp = Z(j4 + 2) / Z(j4 - 3);
            d = d * temp - tau;
            *dmin = fmn(*dmin, d);
            Z(j4 - 1) = Z(j4) * temp;
            emin = fmn(Z(j4 - 1), emin);
        }
    }
    *dnm2 = d;
    *dmin2 = *dmin;
    j4 = 4 * (n0 - 2) - pp;
    j4p2 = j4 + 2 * pp - 1;
    Z(j4 - 2) = *dnm2 + Z(j4p2);
    Z(j4) = Z(j4p2 + 2) * (Z(j4p2) / Z(j4 - 2));
    *dnm1 = Z(j4p2 + 2) * (*dnm2 / Z(j4 - 2)) - tau;
    *dmin = fmn(*dmin, *dnm1);
    *dmin1 = *dmin;
    j4 = j4 + 4;
    j4p2 = j4 + 2 * pp - 1;
    Z(j4 - 2) = *dnm1 + Z(j4p2);
    Z(j4) = Z(j4p2 + 2) * (Z(j4p2) / Z(j4 - 2));
    *dn = Z(j4p2 + 2) * (*dnm1 / Z(j4 - 2)) - tau;
    *dmin = fmn(*dmin, *dn);
    Z(j4 + 2) = *dn;
    Z(4 * n0 - pp) = emin;
}

/* dlasq6: one dqd transform without shift, with underflow guards, odsq6.f. */
static void o_dlasq6(int i0, int n0, double *z, int pp, double *dmin, double *dmin1,
                     double *dmin2, double *dn, double *dnm1, double *dnm2)
{
    int j4, j4p2;
    double d, emin, temp;
    if ((n0 - i0 - 1) <= 0) return;
    j4 = 4 * i0 + pp - 3;
    emin = Z(j4 + 4);
    d = Z(j4);
    *dmin = d;
    if (pp == 0) {
        for (j4 = 4 * i0; j4 <= 4 * (n0 - 3); j4 += 4) {
            Z(j4 - 2) = d + Z(j4 - 1);
            if (Z(j4 - 2) == 0.0) {
                Z(j4) = 0.0; d = Z(j4 + 1); *dmin = d; emin = 0.0;
            } else if (SAFMIN * Z(j4 + 1) < Z(j4 - 2) && SAFMIN * Z(j4 - 2) < Z(j4 + 1)) {
                temp = Z(j4 + 1) / Z(j4 - 2);
                Z(j4) = Z(j4 - 1) * temp;
                d = d * temp;
            } else {
                Z(j4) = Z(j4 + 1) * (Z(j4 - 1) / Z(j4 - 2));
                d = Z(j4 + 1) * (d / Z(j4 - 2));
            }
            *dmin = fmn(*dmin, d);
            emin = fmn(emin, Z(j4));
        }
    } else {
        for (j4 = 4 * i0; j4 <= 4 * (n0 - 3); j4 += 4) {
            Z(j4 - 3) = d + Z(j4);
            if (Z(j4 - 3) == 0.0) {
                Z(j4 - 1) = 0.0; d = Z(j4 + 2); *dmin = d; emin = 0.0;
            } else if (SAFMIN * Z(j4 + 2) < Z(j4 - 3) && SAFMIN * Z(j4 - 3) < Z(j4 + 2)) {
                temp = Z(j4 + 2) / Z(j4 - 3);
                Z(j4 - 1) = Z(j4) * temp;
                d = d * temp;
            } else {
                Z(j4 - 1) = Z(j4 + 2) * (Z(j4) / Z(j4 - 3));
                d = Z(j4 + 2) * (d / Z(j4 - 3));
            }
            *dmin = fmn(*dmin, d);
            emin = fmn(emin, Z(j4 - 1));
        }
    }
    *dnm2 = d;
    *dmin2 = *dmin;
    j4 = 4 * (n0 - 2) - pp;
    j4p2 = j4 + 2 * pp - 1;
    Z(j4 - 2) = *dnm2 + Z(j4p2);
    if (Z(j4 - 2) == 0.0) {
        Z(j4) = 0.0; *dnm1 = Z(j4p2 + 2); *dmin = *dnm1; emin = 0.0;
    } else if (SAFMIN * Z(j4p2 + 2) < Z(j4 - 2) && SAFMIN * Z(j4 - 2) < Z(j4p2 + 2)) {
        temp = Z(j4p2 + 2) / Z(j4 - 2);
        Z(j4) = Z(j4p2) * temp;
        *dnm1 = *dnm2 * temp;
    } else {
        Z(j4) = Z(j4p2 + 2) * (Z(j4p2) / Z(j4 - 2));
        *dnm1 = Z(j4p2 + 2) * (*dnm2 / Z(j4 - 2));
    }
    *dmin = fmn(*dmin, *dnm1);
    *dmin1 = *dmin;
    j4 = j4 + 4;
    j4p2 = j4 + 2 * pp - 1;
    Z(j4 - 2) = *dnm1 + Z(j4p2);
    if (Z(j4 - 2) == 0.0) {
        Z(j4) = 0.0; *dn = Z(j4p2 + 2); *dmin = *dn; emin = 0.0;
    } else if (SAFMIN * Z(j4p2 + 2) < Z(j4 - 2) && SAFMIN * Z(j4 - 2) < Z(j4p2 + 2)) {
        temp = Z(j4p2 + 2) / Z(j4 - 2);
        Z(j4) = Z(j4p2) * temp;
        *dn = *dnm1 * temp;
    } else {
        Z(j4) = Z(j4p2 + 2) * (Z(j4p2) / Z(j4 - 2));
        *dn = Z(j4p2 + 2) * (*dnm1 / Z(j4 - 2));
    }
    *dmin = fmn(*dmin, *dn);
    Z(j4 + 2) = *dn;
    Z(4 * n0 - pp) = emin;
}

typedef struct {
    int n0, pp, nfail, iter, ndiv, ttype;
    double dmin, sigma, desig, qmax, dmin1, dmin2, dn, dn1, dn2, g, tau;
} o_qd;

/* dlasq3: deflation checks, shift and one transform, odsq3.f. */
static void o_dlasq3(int i0, double *z, o_qd *q)
{
    const double CBIAS = 1.50, tol = EPS * 100.0, tol2 = tol * tol;
    int n0 = q->n0, pp = q->pp, n0in = q->n0, ipn4, j4, nn;
    double s, t, temp;
    for (;;) {
        if (n0 < i0) goto done;
        if (n0 == i0) goto l20;
        nn = 4 * n0 + pp;
        if (n0 == i0 + 1) goto l40;
        if (Z(nn - 5) > tol2 * (q->sigma + Z(nn - 3)) && Z(nn - 2 * pp - 4) > tol2 * Z(nn - 7))
            goto l30;
    l20:
        Z(4 * n0 - 3) = Z(4 * n0 + pp - 3) + q->sigma;
        n0 = n0 - 1;
        continue;
    l30:
        if (Z(nn - 9) > tol2 * q->sigma && Z(nn - 2 * pp - 8) > tol2 * Z(nn - 11)) break;
    l40:
        if (Z(nn - 3) > Z(nn - 7)) { s = Z(nn - 3); Z(nn - 3) = Z(nn - 7); Z(nn - 7) = s; }
        if (Z(nn - 5) > Z(nn - 3) * tol2) {
            t = 0.5 * ((Z(nn - 7) - Z(nn - 3)) + Z(nn - 5));
            s = Z(nn - 3) * (Z(nn - 5) / t);
            if (s <= t) s = Z(nn - 3) * (Z(nn - 5) / (t * (1.0 + sqrt(1.0 + s / t))));
            else s = Z(nn - 3) * (Z(nn - 5) / (t + sqrt(t) * sqrt(t + s)));
            t = Z(nn - 7) + (s + Z(nn - 5));
            Z(nn - 3) = Z(nn - 3) * (Z(nn - 7) / t);
            Z(nn - 7) = t;
        }
        Z(4 * n0 - 7) = Z(nn - 7) + q->sigma;
        Z(4 * n0 - 3) = Z(nn - 3) + q->sigma;
        n0 = n0 - 2;
    }
    /* label 50 */
    if (pp == 2) pp = 0;
    if (q->dmin <= 0.0 || n0 < n0in) {
        if (CBIAS * Z(4 * i0 + pp - 3) < Z(4 * n0 + pp - 3)) {
            ipn4 = 4 * (i0 + n0);
            for (j4 = 4 * i0; j4 <= 2 * (i0 + n0 - 1); j4 += 4) {
                temp = Z(j4 - 3); Z(j4 - 3) = Z(ipn4 - j4 - 3); Z(ipn4 - j4 - 3) = temp;
                temp = Z(j4 - 2); Z(j4 - 2) = Z(ipn4 - j4 - 2); Z(ipn4 - j4 - 2) = temp;
                temp = Z(j4 - 1); Z(j4 - 1) = Z(ipn4 - j4 - 5); Z(ipn4 - j4 - 5) = temp;
                temp = Z(j4);     Z(j4) = Z(ipn4 - j4 - 4);     Z(ipn4 - j4 - 4) = temp;
            }
            if (n0 - i0 <= 4) {
                Z(4 * n0 + pp - 1) = Z(4 * i0 + pp - 1);
                Z(4 * n0 - pp) = Z(4 * i0 - pp);
            }
            q->dmin2 = fmn(q->dmin2, Z(4 * n0 + pp - 1));
            Z(4 * n0 + pp - 1) = fmn(fmn(Z(4 * n0 + pp - 1), Z(4 * i0 + pp - 1)), Z(4 * i0 + pp + 3));
            Z(4 * n0 - pp) = fmn(fmn(Z(4 * n0 - pp), Z(4 * i0 - pp)), Z(4 * i0 - pp + 4));
            q->qmax = fmx(fmx(q->qmax, Z(4 * i0 + pp - 3)), Z(4 * i0 + pp + 1));
            q->dmin = -0.0;
        }
    }
    o_dlasq4(i0, n0, z, pp, n0in, q->dmin, q->dmin1, q->dmin2, q->dn, q->dn1, q->dn2,
             &q->tau, &q->ttype, &q->g);
    for (;;) { /* label 70 */
        o_dlasq5(i0, n0, z, pp, q->tau, &q->dmin, &q->dmin1, &q->dmin2, &q->dn, &q->dn1,
                 &q->dn2);
        q->ndiv += n0 - i0 + 2;
        q->iter++;
        if (q->dmin >= 0.0 && q->dmin1 > 0.0) goto l90;
        if (q->dmin < 0.0 && q->dmin1 > 0.0 && Z(4 * (n0 - 1) - pp) < tol * (q->sigma + q->dn1) &&
            fabs(q->dn) < tol * q->sigma) {
            Z(4 * (n0 - 1) - pp + 2) = 0.0;
            q->dmin = 0.0;
            goto l90;
        }
        if (q->dmin < 0.0) {
            q->nfail++;
            if (q->ttype < -22) q->tau = 0.0;
            else if (q->dmin1 > 0.0) {
                q->tau = (q->tau + q->dmin) * (1.0 - 2.0 * EPS);
                q->ttype -= 11;
            } else {
                q->tau = 0.25 * q->tau;
                q->ttype -= 12;
            }
            continue;
        }
        if (isnan(q->dmin)) {
            if (q->tau == 0.0) break;
            q->tau = 0.0;
            continue;
        }
        break;
    }
    /* label 80 */
    o_dlasq6(i0, n0, z, pp, &q->dmin, &q->dmin1, &q->dmin2, &q->dn, &q->dn1, &q->dn2);
    q->ndiv += n0 - i0 + 2;
    q->iter++;
    q->tau = 0.0;
l90:
    if (q->tau < q->sigma) {
        q->desig = q->desig + q->tau;
        t = q->sigma + q->desig;
        q->desig = q->desig - (t - q->sigma);
    } else {
        t = q->sigma + q->tau;
        q->desig = q->sigma - (t - q->tau) + q->desig;
    }
    q->sigma = t;
done:
    q->n0 = n0;
    q->pp = pp;
}

static int cmp_desc(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x > y) ? -1 : (x < y) ? 1 : 0;
}

/* dlasq2: all eigenvalues of the qd array Z(1:2n) (q, e pairs), odsq2.f.
 * On return Z(1:n) holds them in decreasing order.  Returns INFO. */
static int o_dlasq2(int n, double *z)
{
    const double CBIAS = 1.50, tol = EPS * 100.0, tol2 = tol * tol;
    double d, dee, deemin, e, emax, emin, oldemn, qmax, qmin, s, t, temp, trace, zmax;
    double tempe, tempq;
    int i0, i1, i4, ipn4, k, kmin, n1, nbig, splt;
    if (n < 0) return -1;
    if (n == 0) return 0;
    if (n == 1) return (Z(1) < 0.0) ? -201 : 0;
    if (n == 2) {
        if (Z(2) < 0.0 || Z(3) < 0.0) return -2;
        if (Z(3) > Z(1)) { d = Z(3); Z(3) = Z(1); Z(1) = d; }
        Z(5) = Z(1) + Z(2) + Z(3);
        if (Z(2) > Z(3) * tol2) {
            t = 0.5 * ((Z(1) - Z(3)) + Z(2));
            s = Z(3) * (Z(2) / t);
            if (s <= t) s = Z(3) * (Z(2) / (t * (1.0 + sqrt(1.0 + s / t))));
            else s = Z(3) * (Z(2) / (t + sqrt(t) * sqrt(t + s)));
            t = Z(1) + (s + Z(2));
            Z(3) = Z(3) * (Z(1) / t);
            Z(1) = t;
        }
        Z(2) = Z(3);
        Z(6) = Z(2) + Z(1);
        return 0;
    }
    Z(2 * n) = 0.0;
    emin = Z(2);
    qmax = 0.0;
    zmax = 0.0;
    d = 0.0;
    e = 0.0;
    for (k = 1; k <= 2 * (n - 1); k += 2) {
        if (Z(k) < 0.0) return -(200 + k);
        if (Z(k + 1) < 0.0) return -(200 + k + 1);
        d = d + Z(k);
        e = e + Z(k + 1);
        qmax = fmx(qmax, Z(k));
        emin = fmn(emin, Z(k + 1));
        zmax = fmx(fmx(qmax, zmax), Z(k + 1));
    }
    if (Z(2 * n - 1) < 0.0) return -(200 + 2 * n - 1);
    d = d + Z(2 * n - 1);
    qmax = fmx(qmax, Z(2 * n - 1));
    zmax = fmx(qmax, zmax);
    if (e == 0.0) {
        for (k = 2; k <= n; k++) Z(k) = Z(2 * k - 1);
        qsort(z, (size_t)n, sizeof(double), cmp_desc);
        Z(2 * n - 1) = d;
        return 0;
    }
    trace = d + e;
    if (trace == 0.0) { Z(2 * n - 1) = 0.0; return 0; }
    for (k = 2 * n; k >= 2; k -= 2) {
        Z(2 * k) = 0.0;
        Z(2 * k - 1) = Z(k);
        Z(2 * k - 2) = 0.0;
        Z(2 * k - 3) = Z(k - 1);
    }
    i0 = 1;
    int n0 = n;
    if (CBIAS * Z(4 * i0 - 3) < Z(4 * n0 - 3)) {
        ipn4 = 4 * (i0 + n0);
        for (i4 = 4 * i0; i4 <= 2 * (i0 + n0 - 1); i4 += 4) {
            temp = Z(i4 - 3); Z(i4 - 3) = Z(ipn4 - i4 - 3); Z(ipn4 - i4 - 3) = temp;
            temp = Z(i4 - 1); Z(i4 - 1) = Z(ipn4 - i4 - 5); Z(ipn4 - i4 - 5) = temp;
        }
    }
    int pp = 0;
    for (k = 1; k <= 2; k++) {
        d = Z(4 * n0 + pp - 3);
        for (i4 = 4 * (n0 - 1) + pp; i4 >= 4 * i0 + pp; i4 -= 4) {
            if (Z(i4 - 1) <= tol2 * d) { Z(i4 - 1) = -0.0; d = Z(i4 - 3); }
            else d = Z(i4 - 3) * (d / (d + Z(i4 - 1)));
        }
        emin = Z(4 * i0 + pp + 1);
        d = Z(4 * i0 + pp - 3);
        for (i4 = 4 * i0 + pp; i4 <= 4 * (n0 - 1) + pp; i4 += 4) {
            Z(i4 - 2 * pp - 2) = d + Z(i4 - 1);
            if (Z(i4 - 1) <= tol2 * d) {
                Z(i4 - 1) = -0.0;
                Z(i4 - 2 * pp - 2) = d;
                Z(i4 - 2 * pp) = 0.0;
                d = Z(i4 + 1);
            } else if (SAFMIN * Z(i4 + 1) < Z(i4 - 2 * pp - 2) &&
                       SAFMIN * Z(i4 - 2 * pp - 2) < Z(i4 + 1)) {
                temp = Z(i4 + 1) / Z(i4 - 2 * pp - 2);
                Z(i4 - 2 * pp) = Z(i4 - 1) * temp;
                d = d * temp;
            } else {
                Z(i4 - 2 * pp) = Z(i4 + 1) * (Z(i4 - 1) / Z(i4 - 2 * pp - 2));
                d = Z(i4 + 1) * (d / Z(i4 - 2 * pp - 2));
            }
            emin = fmn(emin, Z(i4 - 2 * pp));
        }
        Z(4 * n0 - pp - 2) = d;
        qmax = Z(4 * i0 - pp - 2);
        for (i4 = 4 * i0 - pp + 2; i4 <= 4 * n0 - pp - 2; i4 += 4) qmax = fmx(qmax, Z(i4));
        pp = 1 - pp;
    }
    o_qd q;
    memset(&q, 0, sizeof(q));
    q.ttype = 0;
    q.iter = 2;
    q.nfail = 0;
    q.ndiv = 2 * (n0 - i0);
    q.qmax = qmax;
    for (int iwhila = 1; iwhila <= n + 1; iwhila++) {
        if (n0 < 1) goto l170;
        q.desig = 0.0;
        if (n0 == n) q.sigma = 0.0;
        else q.sigma = -Z(4 * n0 - 1);
        if (q.sigma < 0.0) return 1;
        emax = 0.0;
        if (n0 > i0) emin = fabs(Z(4 * n0 - 5));
        else emin = 0.0;
        qmin = Z(4 * n0 - 3);
        q.qmax = qmin;
        for (i4 = 4 * n0; i4 >= 8; i4 -= 4) {
            if (Z(i4 - 5) <= 0.0) goto l100;
            if (qmin >= 4.0 * emax) {
                qmin = fmn(qmin, Z(i4 - 3));
                emax = fmx(emax, Z(i4 - 5));
            }
            q.qmax = fmx(q.qmax, Z(i4 - 7) + Z(i4 - 5));
            emin = fmn(emin, Z(i4 - 5));
        }
        i4 = 4;
    l100:
        i0 = i4 / 4;
        pp = 0;
        if (n0 - i0 > 1) {
            dee = Z(4 * i0 - 3);
            deemin = dee;
            kmin = i0;
            for (i4 = 4 * i0 + 1; i4 <= 4 * n0 - 3; i4 += 4) {
                dee = Z(i4) * (dee / (dee + Z(i4 - 2)));
                if (dee <= deemin) { deemin = dee; kmin = (i4 + 3) / 4; }
            }
            if ((kmin - i0) * 2 < n0 - kmin && deemin <= 0.5 * Z(4 * n0 - 3)) {
                ipn4 = 4 * (i0 + n0);
                pp = 2;
                for (i4 = 4 * i0; i4 <= 2 * (i0 + n0 - 1); i4 += 4) {
                    temp = Z(i4 - 3); Z(i4 - 3) = Z(ipn4 - i4 - 3); Z(ipn4 - i4 - 3) = temp;
                    temp = Z(i4 - 2); Z(i4 - 2) = Z(ipn4 - i4 - 2); Z(ipn4 - i4 - 2) = temp;
                    temp = Z(i4 - 1); Z(i4 - 1) = Z(ipn4 - i4 - 5); Z(ipn4 - i4 - 5) = temp;
                    temp = Z(i4);     Z(i4) = Z(ipn4 - i4 - 4);     Z(ipn4 - i4 - 4) = temp;
                }
            }
        }
        q.dmin = -fmx(0.0, qmin - 2.0 * sqrt(qmin) * sqrt(emax));
        nbig = 100 * (n0 - i0 + 1);
        for (int iwhilb = 1; iwhilb <= nbig; iwhilb++) {
            if (i0 > n0) goto l150;
            q.n0 = n0;
            q.pp = pp;
            o_dlasq3(i0, z, &q);
            n0 = q.n0;
            pp = q.pp;
            pp = 1 - pp;
            if (pp == 0 && n0 - i0 >= 3) {
                if (Z(4 * n0) <= tol2 * q.qmax || Z(4 * n0 - 1) <= tol2 * q.sigma) {
                    splt = i0 - 1;
                    q.qmax = Z(4 * i0 - 3);
                    emin = Z(4 * i0 - 1);
                    oldemn = Z(4 * i0);
                    for (i4 = 4 * i0; i4 <= 4 * (n0 - 3); i4 += 4) {
                        if (Z(i4) <= tol2 * Z(i4 - 3) || Z(i4 - 1) <= tol2 * q.sigma) {
                            Z(i4 - 1) = -q.sigma;
                            splt = i4 / 4;
                            q.qmax = 0.0;
                            emin = Z(i4 + 3);
                            oldemn = Z(i4 + 4);
                        } else {
                            q.qmax = fmx(q.qmax, Z(i4 + 1));
                            emin = fmn(emin, Z(i4 - 1));
                            oldemn = fmn(oldemn, Z(i4));
                        }
                    }
                    Z(4 * n0 - 1) = emin;
                    Z(4 * n0) = oldemn;
                    i0 = splt + 1;
                }
            }
        }
        /* too many iterations in the inner loop: restore and return (INFO = 2) */
        i1 = i0;
        n1 = n0;
        for (;;) {
            tempq = Z(4 * i0 - 3);
            Z(4 * i0 - 3) = Z(4 * i0 - 3) + q.sigma;
            for (k = i0 + 1; k <= n0; k++) {
                tempe = Z(4 * k - 5);
                Z(4 * k - 5) = Z(4 * k - 5) * (tempq / Z(4 * k - 7));
                tempq = Z(4 * k - 3);
                Z(4 * k - 3) = Z(4 * k - 3) + q.sigma + tempe - Z(4 * k - 5);
            }
            if (i1 > 1) {
                n1 = i1 - 1;
                while (i1 >= 2 && Z(4 * i1 - 5) >= 0.0) i1 = i1 - 1;
                q.sigma = -Z(4 * n1 - 1);
                continue;
            }
            break;
        }
        for (k = 1; k <= n; k++) {
            Z(2 * k - 1) = Z(4 * k - 3);
            Z(2 * k) = (k < n0) ? Z(4 * k - 1) : 0.0;
        }
        return 2;
    l150:;
    }
    return 3;
l170:
    for (k = 2; k <= n; k++) Z(k) = Z(4 * k - 3);
    qsort(z, (size_t)n, sizeof(double), cmp_desc);
    e = 0.0;
    for (k = n; k >= 1; k--) e = e + Z(k);
    Z(2 * n + 1) = trace;
    Z(2 * n + 2) = e;
    Z(2 * n + 3) = (double)q.iter;
    Z(2 * n + 4) = (double)q.ndiv / (double)(n * n);
    Z(2 * n + 5) = 100.0 * q.nfail / (double)q.iter;
    (void)zmax; (void)emin;
    return 0;
}
#undef Z

/* ------------------------------------------------------------------------------
 * dlarre: root representation and eigenvalues of every block, odrre.f:296-891.
 * Arrays 0-based here; isplit, iblock, indexw hold the reference's 1-based values.
 * On return E(iend) of every block holds its shift and D, E its root L D L^T.
 * ---------------------------------------------------------------------------- */
static int o_dlarre(char range, int n, double *vl, double *vu, int il, int iu,
                    double *d, double *e, double *e2, double rtol1, double rtol2,
                    double spltol, int *nsplit, int *isplit, int *m, double *w,
                    double *werr, double *wgap, int *iblock, int *indexw,
                    double *gers, double *pivmin)
{
    const int ALLRNG = 1, INDRNG = 2, VALRNG = 3, MAXTRY = 6;
    const double FAC = 0.5, FUDGE = 2.0, PERT = 8.0, MAXGROWTH = 64.0, HNDRD = 100.0;
    int irange = (range == 'A') ? ALLRNG : (range == 'V') ? VALRNG : INDRNG;
    double rtl = sqrt(EPS), bsrtol = sqrt(EPS);
    *m = 0;
    if (n == 1) {
        if (irange == ALLRNG || (irange == VALRNG && d[0] > *vl && d[0] <= *vu) ||
            (irange == INDRNG && il == 1 && iu == 1)) {
            *m = 1; w[0] = d[0]; werr[0] = 0.0; wgap[0] = 0.0;
            iblock[0] = 1; indexw[0] = 1; gers[0] = d[0]; gers[1] = d[0];
        }
        e[0] = 0.0;
        return 0;
    }
    double gl = d[0], gu = d[0], eold = 0.0, emax = 0.0;
    e[n - 1] = 0.0;
    for (int i = 0; i < n; i++) {
        werr[i] = 0.0;
        wgap[i] = 0.0;
        double eabs = fabs(e[i]);
        if (eabs >= emax) emax = eabs;
        double tmp1 = eabs + eold;
        gers[2 * i] = d[i] - tmp1;
        gl = fmn(gl, gers[2 * i]);
        gers[2 * i + 1] = d[i] + tmp1;
        gu = fmx(gu, gers[2 * i + 1]);
        eold = eabs;
    }
    *pivmin = SAFMIN * fmx(1.0, emax * emax);
    double spdiam = gu - gl;
    o_dlarra(n, d, e, e2, spltol, spdiam, nsplit, isplit);
    const int forceb = 0;
    int usedqd = (irange == ALLRNG && !forceb);
    int mm = 0;
    if (irange == ALLRNG && !forceb) {
        *vl = gl;
        *vu = gu;
    } else {
        double wl_o, wu_o;
        int iinfo = o_dlarrd(range, n, *vl, *vu, il, iu, gers, bsrtol, d, e2, *pivmin,
                             *nsplit, isplit, &mm, w, werr, &wl_o, &wu_o, iblock, indexw);
        if (iinfo != 0) return -1;
        *vl = wl_o;
        *vu = wu_o;
        for (int i = mm; i < n; i++) { w[i] = 0.0; werr[i] = 0.0; iblock[i] = 0; indexw[i] = 0; }
    }
    double *work = (double *)calloc((size_t)(6 * n + 16), sizeof(double));
    int ibegin = 1, wbegin = 1;
    for (int jblk = 1; jblk <= *nsplit; jblk++) {
        int iend = isplit[jblk - 1];
        int in = iend - ibegin + 1;
        if (in == 1) {
            if (irange == ALLRNG || (irange == VALRNG && d[ibegin - 1] > *vl && d[ibegin - 1] <= *vu) ||
                (irange == INDRNG && iblock[wbegin - 1] == jblk)) {
                (*m)++;
                w[*m - 1] = d[ibegin - 1];
                werr[*m - 1] = 0.0;
                wgap[*m - 1] = 0.0;
                iblock[*m - 1] = jblk;
                indexw[*m - 1] = 1;
                wbegin++;
            }
            e[iend - 1] = 0.0;
            ibegin = iend + 1;
            continue;
        }
        e[iend - 1] = 0.0;
        gl = d[ibegin - 1];
        gu = d[ibegin - 1];
        for (int i = ibegin; i <= iend; i++) {
            gl = fmn(gers[2 * i - 2], gl);
            gu = fmx(gers[2 * i - 1], gu);
        }
        spdiam = gu - gl;
        int mb = 0, wend = 0, indl = 0, indu = 0;
        double sigma = 0.0;
        if (!(irange == ALLRNG && !forceb)) {
            for (int i = wbegin; i <= mm; i++) {
                if (iblock[i - 1] == jblk) mb++;
                else break;
            }
            if (mb == 0) {
                e[iend - 1] = 0.0;
                ibegin = iend + 1;
                continue;
            }
            usedqd = (mb > FAC * in && !forceb);
            wend = wbegin + mb - 1;
            sigma = 0.0;
            for (int i = wbegin; i <= wend - 1; i++)
                wgap[i - 1] = fmx(0.0, w[i] - werr[i] - (w[i - 1] + werr[i - 1]));
            wgap[wend - 1] = fmx(0.0, *vu - sigma - (w[wend - 1] + werr[wend - 1]));
            indl = indexw[wbegin - 1];
            indu = indexw[wend - 1];
        }
        double isleft, isrght, tmp, tmp1;
        if ((irange == ALLRNG && !forceb) || usedqd) {
            o_dlarrk(in, 1, gl, gu, d + ibegin - 1, e2 + ibegin - 1, *pivmin, rtl, &tmp, &tmp1);
            isleft = fmx(gl, tmp - tmp1 - HNDRD * EPS * fabs(tmp - tmp1));
            o_dlarrk(in, in, gl, gu, d + ibegin - 1, e2 + ibegin - 1, *pivmin, rtl, &tmp, &tmp1);
            isrght = fmn(gu, tmp + tmp1 + HNDRD * EPS * fabs(tmp + tmp1));
            spdiam = isrght - isleft;
        } else {
            isleft = fmx(gl, w[wbegin - 1] - werr[wbegin - 1] -
                                 HNDRD * EPS * fabs(w[wbegin - 1] - werr[wbegin - 1]));
            isrght = fmn(gu, w[wend - 1] + werr[wend - 1] +
                                 HNDRD * EPS * fabs(w[wend - 1] + werr[wend - 1]));
        }
        double s1, s2;
        if (irange == ALLRNG && !forceb) {
            usedqd = 1;
            indl = 1;
            indu = in;
            mb = in;
            wend = wbegin + mb - 1;
            s1 = isleft + 0.25 * spdiam;
            s2 = isrght - 0.25 * spdiam;
        } else if (usedqd) {
            s1 = isleft + 0.25 * spdiam;
            s2 = isrght - 0.25 * spdiam;
        } else {
            tmp = fmn(isrght, *vu) - fmx(isleft, *vl);
            s1 = fmx(isleft, *vl) + 0.25 * tmp;
            s2 = fmn(isrght, *vu) - 0.25 * tmp;
        }
        int cnt1 = 0, cnt2 = 0;
        if (mb > 1) o_dlarrc_T(in, s1, s2, d + ibegin - 1, e + ibegin - 1, &cnt1, &cnt2);
        double sgndef;
        if (mb == 1) { sigma = gl; sgndef = 1.0; }
        else if (cnt1 - indl >= indu - cnt2) {
            if (irange == ALLRNG && !forceb) sigma = fmx(isleft, gl);
            else if (usedqd) sigma = isleft;
            else sigma = fmx(isleft, *vl);
            sgndef = 1.0;
        } else {
            if (irange == ALLRNG && !forceb) sigma = fmn(isrght, gu);
            else if (usedqd) sigma = isrght;
            else sigma = fmn(isrght, *vu);
            sgndef = -1.0;
        }
        double tau;
        if (usedqd) {
            tau = spdiam * EPS * n + 2.0 * *pivmin;
            tau = fmx(tau, 2.0 * EPS * fabs(sigma));
        } else if (mb > 1) {
            double clwdth = w[wend - 1] + werr[wend - 1] - w[wbegin - 1] - werr[wbegin - 1];
            double avgap = fabs(clwdth / (double)(wend - wbegin));
            if (sgndef == 1.0) {
                tau = 0.5 * fmx(wgap[wbegin - 1], avgap);
                tau = fmx(tau, werr[wbegin - 1]);
            } else {
                tau = 0.5 * fmx(wgap[wend - 2], avgap);
                tau = fmx(tau, werr[wend - 1]);
            }
        } else tau = werr[wbegin - 1];
        int ok = 0;
        for (int idum = 1; idum <= MAXTRY; idum++) {
            double dpivot = d[ibegin - 1] - sigma;
            work[0] = dpivot;
            double dmax = fabs(work[0]);
            int j = ibegin;
            for (int i = 1; i <= in - 1; i++) {
                work[2 * in + i - 1] = 1.0 / work[i - 1];
                tmp = e[j - 1] * work[2 * in + i - 1];
                work[in + i - 1] = tmp;
                dpivot = (d[j] - sigma) - tmp * e[j - 1];
                work[i] = dpivot;
                dmax = fmx(dmax, fabs(dpivot));
                j++;
            }
            int norep = (dmax > MAXGROWTH * spdiam);
            if (usedqd && !norep) {
                for (int i = 1; i <= in; i++) {
                    tmp = sgndef * work[i - 1];
                    if (tmp < 0.0) norep = 1;
                }
            }
            if (norep) {
                if (idum == MAXTRY - 1) {
                    if (sgndef == 1.0) sigma = gl - FUDGE * spdiam * EPS * n - FUDGE * 2.0 * *pivmin;
                    else sigma = gu + FUDGE * spdiam * EPS * n + FUDGE * 2.0 * *pivmin;
                } else {
                    sigma = sigma - sgndef * tau;
                    tau = 2.0 * tau;
                }
            } else { ok = 1; break; }
        }
        if (!ok) { free(work); return 2; }
        e[iend - 1] = sigma;
        memcpy(d + ibegin - 1, work, sizeof(double) * (size_t)in);
        memcpy(e + ibegin - 1, work + in, sizeof(double) * (size_t)(in - 1));
        if (mb > 1) {
            int iseed[4] = {1, 1, 1, 1};
            oracle_dlarnv2(iseed, 2 * in - 1, work);
            for (int i = 1; i <= in - 1; i++) {
                d[ibegin + i - 2] = d[ibegin + i - 2] * (1.0 + EPS * PERT * work[i - 1]);
                e[ibegin + i - 2] = e[ibegin + i - 2] * (1.0 + EPS * PERT * work[in + i - 1]);
            }
            d[iend - 1] = d[iend - 1] * (1.0 + EPS * 4.0 * work[in - 1]);
        }
        if (!usedqd) {
            for (int j = wbegin; j <= wend; j++) {
                w[j - 1] = w[j - 1] - sigma;
                werr[j - 1] = werr[j - 1] + fabs(w[j - 1]) * EPS;
            }
            for (int i = ibegin; i <= iend - 1; i++) work[i - 1] = d[i - 1] * (e[i - 1] * e[i - 1]);
            oracle_dlarrb(in, d + ibegin - 1, work + ibegin - 1, indl, indu, rtol1, rtol2,
                          indl - 1, w + wbegin - 1, wgap + wbegin - 1, werr + wbegin - 1,
                          *pivmin, spdiam, in);
            wgap[wend - 1] = fmx(0.0, (*vu - sigma) - (w[wend - 1] + werr[wend - 1]));
            for (int i = indl; i <= indu; i++) {
                (*m)++;
                iblock[*m - 1] = jblk;
                indexw[*m - 1] = i;
            }
        } else {
            double rtol = log((double)in) * 4.0 * EPS;
            int j = ibegin;
            for (int i = 1; i <= in - 1; i++) {
                work[2 * i - 2] = fabs(d[j - 1]);
                work[2 * i - 1] = e[j - 1] * e[j - 1] * work[2 * i - 2];
                j++;
            }
            work[2 * in - 2] = fabs(d[iend - 1]);
            work[2 * in - 1] = 0.0;
            int iinfo = o_dlasq2(in, work);
            if (iinfo != 0) { free(work); return -5; }
            for (int i = 1; i <= in; i++)
                if (work[i - 1] < 0.0) { free(work); return -6; }
            if (sgndef > 0.0) {
                for (int i = indl; i <= indu; i++) {
                    (*m)++;
                    w[*m - 1] = work[in - i];
                    iblock[*m - 1] = jblk;
                    indexw[*m - 1] = i;
                }
            } else {
                for (int i = indl; i <= indu; i++) {
                    (*m)++;
                    w[*m - 1] = -work[i - 1];
                    iblock[*m - 1] = jblk;
                    indexw[*m - 1] = i;
                }
            }
            for (int i = *m - mb + 1; i <= *m; i++) werr[i - 1] = rtol * fabs(w[i - 1]);
            for (int i = *m - mb + 1; i <= *m - 1; i++)
                wgap[i - 1] = fmx(0.0, w[i] - werr[i] - (w[i - 1] + werr[i - 1]));
            wgap[*m - 1] = fmx(0.0, (*vu - sigma) - (w[*m - 1] + werr[*m - 1]));
        }
        ibegin = iend + 1;
        wbegin = wend + 1;
    }
    free(work);
    return 0;
}

/* ------------------------------------------------------------------------------
 * dlarrv: eigenvectors from the representation tree, odrrv.f:150-896, for the
 * whole spectrum of each block (DOL = 1, DOU = M, as dstemr calls it).
 * d, l: root representations from dlarre (shift of block b in l[isplit(b)-1]).
 * ---------------------------------------------------------------------------- */
static int o_dlarrv(int n, double vl, double *d, double *l, double pivmin,
                    const int *isplit, int m, double minrgp, double rtol1,
                    double rtol2, double *w, double *werr, double *wgap,
                    const int *iblock, const int *indexw, const double *gers,
                    double *z, int ldz, int *isuppz)
{
#define W_(i)      w[(i) - 1]
#define WERR_(i)   werr[(i) - 1]
#define WGAP_(i)   wgap[(i) - 1]
#define WSH(i)     wsh[(i) - 1]
#define D_(i)      d[(i) - 1]
#define L_(i)      l[(i) - 1]
#define ZZ(i, j)   z[(size_t)((i) - 1) + (size_t)((j) - 1) * (size_t)ldz]
#define ISUPPZ(i)  isuppz[(i) - 1]
#define INDEXW(i)  indexw[(i) - 1]
#define IBLOCK(i)  iblock[(i) - 1]
#define TWIST(i)   twist[(i) - 1]
    const int MAXITR = 10;
    double *wsh = (double *)calloc((size_t)(n + 1), sizeof(double));
    double *ld = (double *)calloc((size_t)(n + 1), sizeof(double));
    double *lld = (double *)calloc((size_t)(n + 1), sizeof(double));
    double *wrk = (double *)calloc((size_t)(4 * n + 8), sizeof(double));
    int *twist = (int *)calloc((size_t)(n + 1), sizeof(int));
    int *cl1 = (int *)calloc((size_t)(2 * n + 4), sizeof(int));
    int *cl2 = (int *)calloc((size_t)(2 * n + 4), sizeof(int));
    int info = 0;
    for (int j = 1; j <= m; j++)
        for (int i = 1; i <= n; i++) ZZ(i, j) = 0.0;
    const double rqtol = 2.0 * EPS;
    int ibegin = 1, wbegin = 1;
    for (int jblk = 1; jblk <= IBLOCK(maxi(1, m)); jblk++) {
        int iend = isplit[jblk - 1];
        double sigma = L_(iend);
        int wend = wbegin - 1;
        while (wend < m && IBLOCK(wend + 1) == jblk) wend++;
        if (wend < wbegin) { ibegin = iend + 1; continue; }
        double gl = gers[2 * ibegin - 2], gu = gers[2 * ibegin - 1];
        for (int i = ibegin + 1; i <= iend; i++) {
            gl = fmn(gers[2 * i - 2], gl);
            gu = fmx(gers[2 * i - 1], gu);
        }
        double spdiam = gu - gl;
        int oldien = ibegin - 1, in = iend - ibegin + 1, im = wend - wbegin + 1;
        if (ibegin == iend) {
            ZZ(ibegin, wbegin) = 1.0;
            ISUPPZ(2 * wbegin - 1) = ibegin;
            ISUPPZ(2 * wbegin) = ibegin;
            W_(wbegin) = W_(wbegin) + sigma;
            WSH(wbegin) = W_(wbegin);
            ibegin = iend + 1;
            wbegin = wbegin + 1;
            continue;
        }
        for (int i = 1; i <= im; i++) WSH(wbegin + i - 1) = W_(wbegin + i - 1);
        for (int i = 1; i <= im; i++) W_(wbegin + i - 1) = W_(wbegin + i - 1) + sigma;
        int ndepth = 0, parity = 1, nclus = 1;
        int *oldcl, *newcl;
        cl1[1] = 1;
        cl1[2] = im;
        int idone = 0;
        while (idone < im) {
            if (ndepth > m) { info = -2; goto out; }
            int oldncl = nclus;
            nclus = 0;
            parity = 1 - parity;
            if (parity == 0) { oldcl = cl1; newcl = cl2; }
            else { oldcl = cl2; newcl = cl1; }
            for (int i = 1; i <= oldncl; i++) {
                int oldfst = oldcl[2 * i - 1], oldlst = oldcl[2 * i];
                if (ndepth > 0) {
                    int j = wbegin + oldfst - 1;
                    for (int k = 0; k < in; k++) D_(ibegin + k) = ZZ(ibegin + k, j);
                    for (int k = 0; k < in - 1; k++) L_(ibegin + k) = ZZ(ibegin + k, j + 1);
                    sigma = ZZ(iend, j + 1);
                    for (int k = 0; k < in; k++) { ZZ(ibegin + k, j) = 0.0; ZZ(ibegin + k, j + 1) = 0.0; }
                }
                for (int j = ibegin; j <= iend - 1; j++) {
                    double tmp = D_(j) * L_(j);
                    ld[j] = tmp;
                    lld[j] = tmp * L_(j);
                }
                if (ndepth > 0) {
                    int p = INDEXW(wbegin - 1 + oldfst), q = INDEXW(wbegin - 1 + oldlst);
                    int offset = INDEXW(wbegin) - 1;
                    oracle_dlarrb(in, &D_(ibegin), &lld[ibegin], p, q, rtol1, rtol2, offset,
                                  &WSH(wbegin), &WGAP_(wbegin), &WERR_(wbegin), pivmin, spdiam, in);
                    if (oldfst > 1)
                        WGAP_(wbegin + oldfst - 2) =
                            fmx(WGAP_(wbegin + oldfst - 2),
                                W_(wbegin + oldfst - 1) - WERR_(wbegin + oldfst - 1) -
                                    W_(wbegin + oldfst - 2) - WERR_(wbegin + oldfst - 2));
                    if (wbegin + oldlst - 1 < wend)
                        WGAP_(wbegin + oldlst - 1) =
                            fmx(WGAP_(wbegin + oldlst - 1),
                                W_(wbegin + oldlst) - WERR_(wbegin + oldlst) -
                                    W_(wbegin + oldlst - 1) - WERR_(wbegin + oldlst - 1));
                    for (int j = oldfst; j <= oldlst; j++) W_(wbegin + j - 1) = WSH(wbegin + j - 1) + sigma;
                }
                int newfst = oldfst;
                for (int j = oldfst; j <= oldlst; j++) {
                    int newlst;
                    if (j == oldlst) newlst = j;
                    else if (WGAP_(wbegin + j - 1) >= minrgp * fabs(WSH(wbegin + j - 1))) newlst = j;
                    else continue;
                    int newsiz = newlst - newfst + 1;
                    int newftt = wbegin + newfst - 1;
                    if (newsiz > 1) {
                        double lgap, rgap, tau;
                        if (newfst == 1) lgap = fmx(0.0, W_(wbegin) - WERR_(wbegin) - vl);
                        else lgap = WGAP_(wbegin + newfst - 2);
                        rgap = WGAP_(wbegin + newlst - 1);
                        for (int k = 1; k <= 2; k++) {
                            int p = (k == 1) ? INDEXW(wbegin - 1 + newfst) : INDEXW(wbegin - 1 + newlst);
                            int offset = INDEXW(wbegin) - 1;
                            oracle_dlarrb(in, &D_(ibegin), &lld[ibegin], p, p, rqtol, rqtol, offset,
                                          &WSH(wbegin), &WGAP_(wbegin), &WERR_(wbegin), pivmin,
                                          spdiam, in);
                        }
                        int iinfo = oracle_dlarrf(in, &D_(ibegin), &L_(ibegin), &ld[ibegin], newfst,
                                                  newlst, &WSH(wbegin), &WGAP_(wbegin), &WERR_(wbegin),
                                                  spdiam, lgap, rgap, pivmin, &tau,
                                                  &ZZ(ibegin, newftt), &ZZ(ibegin, newftt + 1), wrk);
                        if (iinfo != 0) { info = -2; goto out; }
                        ZZ(iend, newftt + 1) = sigma + tau;
                        for (int k = newfst; k <= newlst; k++) {
                            double fudge = 3.0 * EPS * fabs(WSH(wbegin + k - 1));
                            WSH(wbegin + k - 1) = WSH(wbegin + k - 1) - tau;
                            fudge = fudge + 4.0 * EPS * fabs(WSH(wbegin + k - 1));
                            WERR_(wbegin + k - 1) = WERR_(wbegin + k - 1) + fudge;
                        }
                        nclus++;
                        newcl[2 * nclus - 1] = newfst;
                        newcl[2 * nclus] = newlst;
                    } else {
                        int iter = 0;
                        double tol = 4.0 * log((double)in) * EPS;
                        int k = newfst;
                        int windex = wbegin + k - 1;
                        int windmn = maxi(windex - 1, 1), windpl = mini(windex + 1, m);
                        double lambda = WSH(windex);
                        double left = WSH(windex) - WERR_(windex);
                        double right = WSH(windex) + WERR_(windex);
                        int indeig = INDEXW(windex);
                        double lgap, rgap, gap, gaptol;
                        if (k == 1) lgap = EPS * fmx(fabs(left), fabs(right));
                        else lgap = WGAP_(windmn);
                        if (k == im) rgap = EPS * fmx(fabs(left), fabs(right));
                        else rgap = WGAP_(windex);
                        gap = fmn(lgap, rgap);
                        if (k == 1 || k == im) gaptol = 0.0;
                        else gaptol = gap * EPS;
                        int isupmn = in, isupmx = 1;
                        double savgap = WGAP_(windex);
                        WGAP_(windex) = gap;
                        int usedbs = 0, usedrq = 0, needbs = 0;
                        int negcnt = 0;
                        double ztz, mingma, nrminv = 0, resid = 0, rqcorr = 0, bstres = 0, bstw = 0;
                        for (;;) { /* label 120 */
                            if (needbs) {
                                usedbs = 1;
                                int itmp1 = TWIST(windex);
                                int offset = INDEXW(wbegin) - 1;
                                oracle_dlarrb(in, &D_(ibegin), &lld[ibegin], indeig, indeig, 0.0,
                                              2.0 * EPS, offset, &WSH(wbegin), &WGAP_(wbegin),
                                              &WERR_(wbegin), pivmin, spdiam, itmp1);
                                lambda = WSH(windex);
                                TWIST(windex) = 0;
                            }
                            oracle_dlar1v(in, 1, in, lambda, &D_(ibegin), &L_(ibegin), &ld[ibegin],
                                          &lld[ibegin], pivmin, gaptol, &ZZ(ibegin, windex), !usedbs,
                                          &negcnt, &ztz, &mingma, &TWIST(windex),
                                          &ISUPPZ(2 * windex - 1), &nrminv, &resid, &rqcorr, wrk);
                            if (iter == 0) { bstres = resid; bstw = lambda; }
                            else if (resid < bstres) { bstres = resid; bstw = lambda; }
                            isupmn = mini(isupmn, ISUPPZ(2 * windex - 1));
                            isupmx = maxi(isupmx, ISUPPZ(2 * windex));
                            iter++;
                            if (resid > tol * gap && fabs(rqcorr) > rqtol * fabs(lambda) && !usedbs) {
                                double sgndef = (indeig <= negcnt) ? -1.0 : 1.0;
                                if (rqcorr * sgndef >= 0.0 && lambda + rqcorr <= right &&
                                    lambda + rqcorr >= left) {
                                    usedrq = 1;
                                    if (sgndef == 1.0) left = lambda;
                                    else right = lambda;
                                    WSH(windex) = 0.5 * (right + left);
                                    lambda = lambda + rqcorr;
                                    WERR_(windex) = 0.5 * (right - left);
                                } else needbs = 1;
                                if (right - left < rqtol * fabs(lambda)) { usedbs = 1; continue; }
                                else if (iter < MAXITR) continue;
                                else if (iter == MAXITR) { needbs = 1; continue; }
                                else { info = 5; goto out; }
                            } else {
                                if (usedrq && usedbs && bstres <= resid) {
                                    lambda = bstw;
                                    oracle_dlar1v(in, 1, in, lambda, &D_(ibegin), &L_(ibegin),
                                                  &ld[ibegin], &lld[ibegin], pivmin, gaptol,
                                                  &ZZ(ibegin, windex), !usedbs, &negcnt, &ztz, &mingma,
                                                  &TWIST(windex), &ISUPPZ(2 * windex - 1), &nrminv,
                                                  &resid, &rqcorr, wrk);
                                }
                                WSH(windex) = lambda;
                                break;
                            }
                        }
                        ISUPPZ(2 * windex - 1) += oldien;
                        ISUPPZ(2 * windex) += oldien;
                        int zfrom = ISUPPZ(2 * windex - 1), zto = ISUPPZ(2 * windex);
                        isupmn += oldien;
                        isupmx += oldien;
                        for (int ii = isupmn; ii <= zfrom - 1; ii++) ZZ(ii, windex) = 0.0;
                        for (int ii = zto + 1; ii <= isupmx; ii++) ZZ(ii, windex) = 0.0;
                        for (int ii = zfrom; ii <= zto; ii++) ZZ(ii, windex) = nrminv * ZZ(ii, windex);
                        W_(windex) = lambda + sigma;
                        if (k > 1)
                            WGAP_(windmn) = fmx(WGAP_(windmn), W_(windex) - WERR_(windex) -
                                                                   W_(windmn) - WERR_(windmn));
                        if (windex < wend)
                            WGAP_(windex) = fmx(savgap, W_(windpl) - WERR_(windpl) -
                                                            W_(windex) - WERR_(windex));
                        idone++;
                    }
                    newfst = j + 1;
                }
            }
            ndepth++;
        }
        ibegin = iend + 1;
        wbegin = wend + 1;
    }
out:
    free(wsh); free(ld); free(lld); free(wrk); free(twist); free(cl1); free(cl2);
    return info;
#undef W_
#undef WERR_
#undef WGAP_
#undef WSH
#undef D_
#undef L_
#undef ZZ
#undef ISUPPZ
#undef INDEXW
#undef IBLOCK
#undef TWIST
}

static int cmp_asc(const void *a, const void *b)
{
    double x = *(const double *)a, y = *(const double *)b;
    return (x < y) ? -1 : (x > y) ? 1 : 0;
}

/* ------------------------------------------------------------------------------
 * dstemr, odstmr.f:233-647, for the calls mrrr() makes (NZC = n, or the count
 * query NZC = -1 of range V).  Returns dstemr's INFO.
 * ---------------------------------------------------------------------------- */
int oracle_dstemr(char jobz, char range, int n, double *D, double *E, double vl,
                  double vu, int il, int iu, int *m, double *W, double *Z, int ldz,
                  int nzc, int *isuppz, int *tryrac)
{
    int wantz = (jobz == 'V' || jobz == 'v');
    int alleig = (range == 'A' || range == 'a');
    int valeig = (range == 'V' || range == 'v');
    int indeig = (range == 'I' || range == 'i');
    int zquery = (nzc == -1);
    double wl = 0.0, wu = 0.0;
    int iil = 0, iiu = 0;
    if (valeig) { wl = vl; wu = vu; }
    else if (indeig) { iil = il; iiu = iu; }
    if (!(wantz || jobz == 'N' || jobz == 'n')) return -1;
    if (!(alleig || valeig || indeig)) return -2;
    if (n < 0) return -3;
    if (valeig && n > 0 && wu <= wl) return -7;
    if (indeig && (iil < 1 || iil > n)) return -8;
    if (indeig && (iiu < iil || iiu > n)) return -9;
    if (ldz < 1 || (wantz && ldz < n)) return -13;
    double smlnum = SAFMIN / EPS, bignum = 1.0 / smlnum;
    double rmin = sqrt(smlnum), rmax = fmn(sqrt(bignum), 1.0 / sqrt(sqrt(SAFMIN)));
    if (wantz && zquery) {
        int nzcmin = 0;
        if (alleig) nzcmin = n;
        else if (valeig) {
            int lc, rc;
            o_dlarrc_T(n, vl, vu, D, E, &lc, &rc);
            nzcmin = rc - lc;
        } else nzcmin = iiu - iil + 1;
        Z[0] = (double)nzcmin;
        return 0;
    }
    *m = 0;
    if (n == 0) return 0;
    if (n == 1) {
        if (alleig || indeig) { *m = 1; W[0] = D[0]; }
        else if (wl < D[0] && wu >= D[0]) { *m = 1; W[0] = D[0]; }
        if (wantz) { Z[0] = 1.0; isuppz[0] = 1; isuppz[1] = 1; }
        return 0;
    }
    if (n == 2) {
        double r1, r2, cs = 0.0, sn = 0.0;
        o_dlaev2(D[0], E[0], D[1], &r1, &r2, &cs, &sn);
        for (int pass = 0; pass < 2; pass++) {
            double r = pass == 0 ? r2 : r1;
            int take = alleig || (valeig && r > wl && r <= wu) ||
                       (indeig && (pass == 0 ? iil == 1 : iiu == 2));
            if (!take) continue;
            (*m)++;
            W[*m - 1] = r;
            if (wantz) {
                double z1 = pass == 0 ? -sn : cs, z2 = pass == 0 ? cs : sn;
                Z[(size_t)(*m - 1) * ldz] = z1;
                Z[(size_t)(*m - 1) * ldz + 1] = z2;
                int a, b;
                if (sn != 0.0) { if (cs != 0.0) { a = 1; b = 2; } else { a = 1; b = 1; } }
                else { a = 2; b = 2; }
                isuppz[2 * *m - 2] = a;
                isuppz[2 * *m - 1] = b;
            }
        }
        return 0;
    }
    /* odnst 'M': largest absolute entry, NaN propagating */
    double tnrm = 0.0;
    for (int i = 0; i < n; i++) { double a = fabs(D[i]); if (tnrm < a || isnan(a)) tnrm = a; }
    for (int i = 0; i < n - 1; i++) { double a = fabs(E[i]); if (tnrm < a || isnan(a)) tnrm = a; }
    double scale = 1.0;
    if (tnrm > 0.0 && tnrm < rmin) scale = rmin / tnrm;
    else if (tnrm > rmax) scale = rmax / tnrm;
    if (scale != 1.0) {
        for (int i = 0; i < n; i++) D[i] = scale * D[i];
        for (int i = 0; i < n - 1; i++) E[i] = scale * E[i];
        tnrm = tnrm * scale;
        if (valeig) { wl = wl * scale; wu = wu * scale; }
    }
    int iinfo = (*tryrac) ? o_dlarrr(n, D, E) : -1;
    double thresh;
    if (iinfo == 0) thresh = EPS;
    else { thresh = -EPS; *tryrac = 0; }
    double *dcopy = NULL;
    if (*tryrac) {
        dcopy = (double *)malloc(sizeof(double) * (size_t)n);
        memcpy(dcopy, D, sizeof(double) * (size_t)n);
    }
    double *e2 = (double *)calloc((size_t)n, sizeof(double));
    for (int j = 0; j < n - 1; j++) e2[j] = E[j] * E[j];
    double rtol1, rtol2;
    if (!wantz) { rtol1 = 4.0 * EPS; rtol2 = 4.0 * EPS; }
    else { rtol1 = sqrt(EPS); rtol2 = fmx(sqrt(EPS) * 5.0e-3, 4.0 * EPS); }
    double *werr = (double *)calloc((size_t)n, sizeof(double));
    double *wgap = (double *)calloc((size_t)n, sizeof(double));
    double *gers = (double *)calloc((size_t)(2 * n), sizeof(double));
    int *isplit = (int *)calloc((size_t)n, sizeof(int));
    int *iblock = (int *)calloc((size_t)n, sizeof(int));
    int *indexw = (int *)calloc((size_t)n, sizeof(int));
    int nsplit = 0, info = 0;
    double pivmin = 0.0;
    iinfo = o_dlarre(range, n, &wl, &wu, iil, iiu, D, E, e2, rtol1, rtol2, thresh, &nsplit,
                     isplit, m, W, werr, wgap, iblock, indexw, gers, &pivmin);
    if (iinfo != 0) { info = 10 + abs(iinfo); goto out; }
    if (wantz) {
        iinfo = o_dlarrv(n, wl, D, E, pivmin, isplit, *m, 1.0e-3, rtol1, rtol2, W, werr, wgap,
                         iblock, indexw, gers, Z, ldz, isuppz);
        if (iinfo != 0) { info = 20 + abs(iinfo); goto out; }
    } else {
        for (int j = 0; j < *m; j++) W[j] = W[j] + E[isplit[iblock[j] - 1] - 1];
    }
    if (*tryrac) {
        int ibegin = 1, wbegin = 1;
        for (int jblk = 1; jblk <= iblock[*m - 1]; jblk++) {
            int iend = isplit[jblk - 1], in = iend - ibegin + 1;
            int wend = wbegin - 1;
            while (wend < *m && iblock[wend] == jblk) wend++;
            if (wend < wbegin) { ibegin = iend + 1; continue; }
            int offset = indexw[wbegin - 1] - 1;
            int ifirst = indexw[wbegin - 1], ilast = indexw[wend - 1];
            o_dlarrj(in, dcopy + ibegin - 1, e2 + ibegin - 1, ifirst, ilast, 4.0 * EPS, offset,
                     W + wbegin - 1, werr + wbegin - 1, pivmin, tnrm);
            ibegin = iend + 1;
            wbegin = wend + 1;
        }
    }
    if (scale != 1.0) {
        double inv = 1.0 / scale;
        for (int j = 0; j < *m; j++) W[j] = inv * W[j];
    }
    if (nsplit > 1) {
        if (!wantz) qsort(W, (size_t)*m, sizeof(double), cmp_asc);
        else {
            for (int j = 1; j <= *m - 1; j++) {
                int i = 0;
                double tmp = W[j - 1];
                for (int jj = j + 1; jj <= *m; jj++)
                    if (W[jj - 1] < tmp) { i = jj; tmp = W[jj - 1]; }
                if (i != 0) {
                    W[i - 1] = W[j - 1];
                    W[j - 1] = tmp;
                    for (int r = 0; r < n; r++) {
                        double t = Z[(size_t)(i - 1) * ldz + r];
                        Z[(size_t)(i - 1) * ldz + r] = Z[(size_t)(j - 1) * ldz + r];
                        Z[(size_t)(j - 1) * ldz + r] = t;
                    }
                    int t1 = isuppz[2 * i - 2]; isuppz[2 * i - 2] = isuppz[2 * j - 2]; isuppz[2 * j - 2] = t1;
                    int t2 = isuppz[2 * i - 1]; isuppz[2 * i - 1] = isuppz[2 * j - 1]; isuppz[2 * j - 1] = t2;
                }
            }
        }
    }
out:
    free(dcopy); free(e2); free(werr); free(wgap); free(gers); free(isplit); free(iblock); free(indexw);
    return info;
}
