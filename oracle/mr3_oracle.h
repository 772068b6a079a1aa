/*
 * mr3_oracle.h -- CPU oracle for the MRRR tridiagonal eigensolver.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a sequential C restatement of the reference
 * algorithm (HPAC/mr3smp: SRC/mrrr.c, SRC/mrrr_val.c, SRC/mrrr_vec.c,
 * SRC/process_{cluster,singleton}.c and the vendored LAPACK kernels under
 * SRC/LAPACK/od*.f).  It is used by the tests, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg as the checker, never by the product path
 * (mr3smp_amd/ never links, loads or calls it).
 *
 * Scope of the restatement (the routing of mrrr.c:141-153 with PMR_NUM_THREADS >= 13,
 * so that the task-parallel route takes its bisection branch for range A):
 *   - n < 500: LAPACK dstemr as vendored by the reference (mr3_oracle_dstemr.c:
 *     odstmr.f, odrre.f with dqds odsq2-6.f, odrrk.f, odrrv.f), bit for bit,
 *     n = 1, 2 closed forms included;
 *   - n >= 500: the task-parallel mrrr() path (mrrr.c:155-296) run on one thread,
 *     with root eigenvalues by bisection (FORCE_BISECTION, mrrr.h:160-162; the dqds
 *     branch mrrr_val.c:464-507 is taken by the reference only with <= 12 threads).
 *     Its task sizes follow one thread; the reference's depend on its thread count
 *     and timing, so results there agree with the reference to tolerance.
 */
#ifndef MR3_ORACLE_H
#define MR3_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Same prototype and argument meaning as the reference mrrr() (INCLUDE/mrrr.h:62-65). */
int oracle_mrrr(char *jobz, char *range, int *n, double *D, double *E,
                double *vl, double *vu, int *il, int *iu, int *tryrac,
                int *m, double *W, double *Z, int *ldz, int *Zsupp);

/* mrrr() sends every matrix with n below this to dstemr (mrrr.c:149-153;
 * DSTEMR_IF_SMALLER, mrrr.h:158). */
#define MR3_DSTEMR_BELOW 500

/* LAPACK dstemr as vendored by the reference (odstmr.f:233-647), for the calls that
 * mrrr() makes: NZC = n, or NZC = -1 (count query, the count lands in Z[0]).
 * Arrays 0-based; isuppz 1-based pairs.  Returns dstemr's INFO. */
int oracle_dstemr(char jobz, char range, int n, double *D, double *E, double vl,
                  double vu, int il, int iu, int *m, double *W, double *Z, int ldz,
                  int nzc, int *isuppz, int *tryrac);

/* Kernel-level entry points (1-based index arguments as in the Fortran). */
/* singletons that fell back to bisection since the last reset (a test statistic) */
long long oracle_fallbacks(int reset);
int    oracle_dlaneg(int n, const double *d, const double *lld, double sigma,
                     double pivmin, int r);                         /* odneg.f  */
int    oracle_sturm_count_T(int n, const double *d, const double *e2,
                            double x, double pivmin);               /* odebz.f:265-277 (IJOB=1) */
void   oracle_dlarnv2(int iseed[4], int n, double *x);              /* odrnv.f, IDIST=2 */
void   oracle_dlar1v(int n, int b1, int bn, double lambda, const double *d,
                     const double *l, const double *ld, const double *lld,
                     double pivmin, double gaptol, double *z, int wantnc,
                     int *negcnt, double *ztz, double *mingma, int *r,
                     int *isuppz, double *nrminv, double *resid,
                     double *rqcorr, double *work);                 /* odr1v.f  */
int    oracle_dlarrf(int n, const double *d, const double *l, const double *ld,
                     int clstrt, int clend, const double *w, const double *wgap,
                     const double *werr, double spdiam, double clgapl,
                     double clgapr, double pivmin, double *sigma,
                     double *dplus, double *lplus, double *work);   /* odrrf.f  */
void   oracle_dlarrb(int n, const double *d, const double *lld, int ifirst,
                     int ilast, double rtol1, double rtol2, int offset,
                     double *w, double *wgap, double *werr, double pivmin,
                     double spdiam, int twist);                     /* odrrb.f  */

#ifdef __cplusplus
}
#endif
#endif
