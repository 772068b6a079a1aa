/*
 * mr3_oracle_int.h -- helpers shared between the two files of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY (see mr3_oracle.h).  The functions are restatements of
 * the reference's vendored LAPACK kernels; each definition cites its source.
 */
#ifndef MR3_ORACLE_INT_H
#define MR3_ORACLE_INT_H

/* dlarrd, ORDER = 'B' (odrrd.f:241-722) */
int  o_dlarrd(char range, int n, double vl, double vu, int il, int iu,
              const double *gers, double reltol, const double *d,
              const double *e2, double pivmin, int nsplit,
              const int *isplit, int *m, double *w, double *werr,
              double *wl_out, double *wu_out, int *iblock, int *indexw);
/* dlarrc, JOBT = 'T' (odrrc.f:95-113) */
void o_dlarrc_T(int n, double vl, double vu, const double *d,
                const double *e, int *lcnt, int *rcnt);
/* dlarra (odrra.f:90-123) */
void o_dlarra(int n, const double *d, double *e, double *e2, double spltol,
              double tnrm, int *nsplit, int *isplit);
/* dlarrr (odrrr.f:60-121): 0 when relative accuracy is attainable */
int  o_dlarrr(int n, const double *d, const double *e);
/* dlarrj (odrrj.f:100-281) */
void o_dlarrj(int n, const double *d, const double *e2, int ifirst,
              int ilast, double rtol, int offset, double *w,
              double *werr, double pivmin, double spdiam);
/* dlaev2 (odev2.f:70-170) */
void o_dlaev2(double a, double b, double c, double *rt1, double *rt2,
              double *cs1, double *sn1);

#endif
